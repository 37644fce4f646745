// knp/backends/gpu-mi355x/backend.h - the MI355X backend, a drop-in sibling of SingleThreadedCPUBackend
// (knp/backends/cpu/cpu-single-threaded-backend/include/knp/backends/cpu-single-threaded/backend.h:52-453):
// same create(), load_populations/load_projections, start/stop_learning, _init and _step, same
// SpikeMessage traffic on the message bus. The network state lives on the GPU; population spikes are sent
// on the bus only for populations some other endpoint subscribes to. The library exports the
// create_knp_backend alias that BackendLoader resolves (backend.cpp:340 of the reference ST backend);
// libknp-mi355x-mt-backend.so exports one that creates the backend with the multi-threaded semantics.
#pragma once
#include <memory>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "../../core/backend.h"

struct knp_net;

namespace knp::backends::gpu_mi355x {

class MI355XBackend : public core::Backend {
public:
    using PopulationVariants = core::AllPopulationsVariant;
    using ProjectionVariants = core::AllProjectionsVariant;

    // Whose projection messages the device network follows (knp_mi355x_set_message_semantics):
    // SingleThreadedCPUBackend's, or MultiThreadedCPUBackend's (cpu-multi-threaded-backend/impl/backend.cpp:
    // 250-295), whose supported types the backend then reports (its backend.h:80-88).
    enum class Semantics { single_threaded, multi_threaded };

    static std::shared_ptr<MI355XBackend> create();
    explicit MI355XBackend(int device = 0, Semantics semantics = Semantics::single_threaded);
    ~MI355XBackend() override;
    MI355XBackend(const MI355XBackend &) = delete;
    MI355XBackend &operator=(const MI355XBackend &) = delete;

    bool plasticity_supported() const override { return true; }
    std::vector<std::string> get_supported_neurons() const override;
    std::vector<std::string> get_supported_synapses() const override;
    // every type of AllPopulationsVariant / AllProjectionsVariant
    std::vector<size_t> get_supported_population_indexes() const override;
    std::vector<size_t> get_supported_projection_indexes() const override;

    void load_populations(const std::vector<PopulationVariants> &populations);
    void load_projections(const std::vector<ProjectionVariants> &projections);
    void load_all_populations(const std::vector<core::AllPopulationsVariant> &p) override { load_populations(p); }
    void load_all_projections(const std::vector<core::AllProjectionsVariant> &p) override { load_projections(p); }
    void remove_projections(const std::vector<core::UID> &uids) override;
    void remove_populations(const std::vector<core::UID> &uids) override;

    std::vector<std::unique_ptr<core::Device>> get_devices() const override;
    // Selecting a device other than the one the backend was created for recreates the device network; that is
    // allowed until the first step.
    void select_devices(const std::set<core::UID> &uids) override;
    void select_device(std::unique_ptr<core::Device> &&device) override;

    void stop_learning() override;
    void start_learning() override;
    void _init() override;
    void _step() override;

    // Backend::get_network_data (knp/core-library/include/knp/core/backend.h:396-420): ranges over copies of
    // the loaded populations and projections carrying the device state at the time of the call.
    DataRanges get_network_data() const override;
    // The same copies as two vectors.
    std::pair<std::vector<PopulationVariants>, std::vector<ProjectionVariants>> network_copy() const;

    knp_net *native_handle() { return net_; }

private:
    void check(int rc, const char *what) const;
    void use_device(const core::UID &uid);
    int device_ = 0;
    Semantics semantics_ = Semantics::single_threaded;
    knp_net *net_ = nullptr;
    bool learning_ = false;
    std::vector<PopulationVariants> populations_;
    std::vector<ProjectionVariants> projections_;
    std::vector<core::UID> pop_uids_, proj_uids_;
};

}  // namespace knp::backends::gpu_mi355x
