// libknp-mi355x-mt-backend.so: the MI355X backend as a drop-in beside MultiThreadedCPUBackend: the same device
// network with that backend's projection messages (the first SpikeMessage of a step only, repeated indexes as
// weight multiples; cpu-multi-threaded-backend/impl/backend.cpp:250-295) and supported types.
#include "knp/backends/gpu-mi355x/backend.h"

namespace {
std::shared_ptr<knp::core::Backend> create_backend() {
    using knp::backends::gpu_mi355x::MI355XBackend;
    return std::make_shared<MI355XBackend>(0, MI355XBackend::Semantics::multi_threaded);
}
}  // namespace

// The factory alias BackendLoader resolves (BOOST_DLL_ALIAS(..., create_knp_backend),
// knp/backends/cpu/cpu-multi-threaded-backend/impl/backend.cpp:447).
extern "C" {
__attribute__((visibility("default"))) std::shared_ptr<knp::core::Backend> (*create_knp_backend)() = &create_backend;
}
