// knp/synapse-traits/output_types.h - interface of knp/synapse-traits-library/.../output_types.h:36-62.
#pragma once
namespace knp::synapse_traits {
enum class OutputType { EXCITATORY = 0, INHIBITORY_CURRENT = 1, INHIBITORY_CONDUCTANCE = 2, DOPAMINE = 3, BLOCKING = 4 };
}  // namespace knp::synapse_traits
