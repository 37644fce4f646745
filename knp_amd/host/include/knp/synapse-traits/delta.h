// knp/synapse-traits/delta.h - DeltaSynapse and its STDP variants.
// Interfaces of knp/synapse-traits-library/include/knp/synapse-traits/{delta.h:46-93, stdp_common.h:42-90,
// stdp_synaptic_resource_rule.h:43-100, stdp_add_rule.h:40-67, stdp_type_traits.h}: same names, fields and
// defaults.
#pragma once
#include <cstdint>
#include <vector>

#include "output_types.h"

namespace knp::synapse_traits {

template <class SynapseType>
struct synapse_parameters;
template <class SynapseType>
struct shared_synapse_parameters {};

struct DeltaSynapse;

template <>
struct synapse_parameters<DeltaSynapse> {
    float weight_ = 0.0F;
    uint32_t delay_ = 1;
    OutputType output_type_ = OutputType::EXCITATORY;
};

template <template <typename> typename Rule, typename Synapse>
struct STDP;

template <typename SynapseType>
struct STDPSynapticResourceRule {
    using LinkedSynapseType = SynapseType;
    float synaptic_resource_ = 0;
    float w_min_ = 0;
    float w_max_ = 1;
    float d_u_ = 0;
    uint32_t dopamine_plasticity_period_ = 0;
    uint64_t last_spike_step_ = 0;
    bool had_hebbian_update_ = false;
    bool has_contributed_ = false;
};

template <typename SynapseType>
struct STDPAdditiveRule {
    using LinkedSynapseType = SynapseType;
    float tau_plus_ = 10;
    float tau_minus_ = 10;
    std::vector<uint64_t> presynaptic_spike_times_;
    std::vector<uint64_t> postsynaptic_spike_times_;
};

template <template <typename> typename Rule, typename Synapse>
struct synapse_parameters<STDP<Rule, Synapse>> : public synapse_parameters<Synapse> {
    using RuleType = Rule<Synapse>;
    using SynapseType = Synapse;
    RuleType rule_;
};

using SynapticResourceSTDPDeltaSynapse = STDP<STDPSynapticResourceRule, DeltaSynapse>;
using AdditiveSTDPDeltaSynapse = STDP<STDPAdditiveRule, DeltaSynapse>;

}  // namespace knp::synapse_traits
