// knp/framework/backend_loader.h - loads a backend shared library by path.
// Behaviour of knp/base-framework/impl/backend_loader.cpp:34-77 (boost::dll::import_alias of the
// `create_knp_backend` alias; is_backend = the alias is exported; a missing library throws), with dlopen.
// The alias is an exported data symbol holding the address of a factory
// `std::shared_ptr<knp::core::Backend>()`, as BOOST_DLL_ALIAS emits it.
#pragma once
#include <dlfcn.h>

#include <filesystem>
#include <memory>
#include <stdexcept>
#include <string>

#include "../core/backend.h"

namespace knp::framework {

class BackendLoader {
public:
    using BackendCreateFunction = std::shared_ptr<core::Backend>();

    std::shared_ptr<core::Backend> load(const std::filesystem::path &backend_path) {
        void *h = open(backend_path);
        void *sym = dlsym(h, "create_knp_backend");
        if (!sym) throw std::runtime_error("Backend library has no create_knp_backend: " + backend_path.string());
        auto create = *reinterpret_cast<BackendCreateFunction **>(sym);
        return create();
    }

    static bool is_backend(const std::filesystem::path &backend_path) {
        void *h = dlopen(decorate(backend_path).c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!h) return false;
        const bool ok = dlsym(h, "create_knp_backend") != nullptr;
        dlclose(h);
        return ok;
    }

private:
    // boost::dll::load_mode::append_decorations: "dir/knp-foo" -> "dir/libknp-foo.so"
    static std::string decorate(const std::filesystem::path &p) {
        if (p.extension() == ".so") return p.string();
        return (p.parent_path() / ("lib" + p.filename().string() + ".so")).string();
    }
    void *open(const std::filesystem::path &p) {
        const std::string name = decorate(p);
        void *h = dlopen(name.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!h) throw std::runtime_error("Cannot load backend " + name + ": " + dlerror());
        return h;  // kept loaded for the process lifetime (backends created from it may outlive the loader)
    }
};

}  // namespace knp::framework
