// libknp-mi355x-backend.so: the MI355X backend as a drop-in beside SingleThreadedCPUBackend.
#include "knp/backends/gpu-mi355x/backend.h"

namespace {
std::shared_ptr<knp::core::Backend> create_backend() {
    return std::make_shared<knp::backends::gpu_mi355x::MI355XBackend>();
}
}  // namespace

// The factory alias BackendLoader resolves (BOOST_DLL_ALIAS(..., create_knp_backend) in the reference,
// knp/backends/cpu/cpu-single-threaded-backend/impl/backend.cpp:340).
extern "C" {
__attribute__((visibility("default"))) std::shared_ptr<knp::core::Backend> (*create_knp_backend)() = &create_backend;
}
