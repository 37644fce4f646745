// knp/core/projection.h - a projection of synapses of one type between two populations.
// Interface of knp/core-library/include/knp/core/projection.h:75-468: Synapse = (parameters, presynaptic
// index, postsynaptic index), construction from a generator, lock state (locked by default, :432), the
// STDP shared parameters with stdp_populations_ (:139-158), and find_synapses (ascending synapse index).
#pragma once
#include <functional>
#include <optional>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "uid.h"
#include "../synapse-traits/delta.h"

namespace knp::core {

enum SynapseElementAccess { synapse_data = 0, source_neuron_id = 1, target_neuron_id = 2 };

template <class SynapseType>
struct SharedSynapseParametersT {
    knp::synapse_traits::shared_synapse_parameters<SynapseType> synapses_parameters_;
};

template <template <typename> typename Rule, typename SynapseT>
struct SharedSynapseParametersT<knp::synapse_traits::STDP<Rule, SynapseT>> {
    enum class ProcessingType { STDPOnly, STDPAndSpike };
    using ContainerType = std::unordered_map<UID, ProcessingType, uid_hash>;
    ContainerType stdp_populations_;
    knp::synapse_traits::shared_synapse_parameters<knp::synapse_traits::STDP<Rule, SynapseT>> synapses_parameters_;
};

template <class SynapseType>
class Projection final {
public:
    using ProjectionSynapseType = SynapseType;
    using SynapseParameters = synapse_traits::synapse_parameters<SynapseType>;
    using Synapse = std::tuple<SynapseParameters, size_t, size_t>;
    using SynapseGenerator = std::function<std::optional<Synapse>(size_t)>;
    using SharedSynapseParameters = SharedSynapseParametersT<SynapseType>;
    enum class Search { by_presynaptic, by_postsynaptic };

    Projection(UID presynaptic_uid, UID postsynaptic_uid) : Projection(UID(), presynaptic_uid, postsynaptic_uid) {}
    Projection(UID uid, UID presynaptic_uid, UID postsynaptic_uid)
        : uid_(uid), presynaptic_uid_(presynaptic_uid), postsynaptic_uid_(postsynaptic_uid) {}
    Projection(UID presynaptic_uid, UID postsynaptic_uid, SynapseGenerator generator, size_t num_iterations)
        : Projection(UID(), presynaptic_uid, postsynaptic_uid, std::move(generator), num_iterations) {}
    Projection(UID uid, UID presynaptic_uid, UID postsynaptic_uid, SynapseGenerator generator, size_t num_iterations)
        : Projection(uid, presynaptic_uid, postsynaptic_uid) {
        add_synapses(std::move(generator), num_iterations);
    }

    const UID &get_uid() const { return uid_; }
    const UID &get_presynaptic() const { return presynaptic_uid_; }
    const UID &get_postsynaptic() const { return postsynaptic_uid_; }
    size_t size() const { return synapses_.size(); }
    Synapse &operator[](size_t i) { return synapses_[i]; }
    const Synapse &operator[](size_t i) const { return synapses_[i]; }
    auto begin() { return synapses_.begin(); }
    auto end() { return synapses_.end(); }
    auto begin() const { return synapses_.cbegin(); }
    auto end() const { return synapses_.cend(); }

    size_t add_synapses(SynapseGenerator generator, size_t num_iterations) {
        size_t added = 0;
        for (size_t i = 0; i < num_iterations; ++i) {
            auto s = generator(i);
            if (s) {
                synapses_.push_back(*s);
                ++added;
            }
        }
        return added;
    }

    std::vector<size_t> find_synapses(size_t neuron_index, Search method) const {
        std::vector<size_t> out;
        for (size_t i = 0; i < synapses_.size(); ++i) {
            size_t v = method == Search::by_presynaptic ? std::get<source_neuron_id>(synapses_[i])
                                                        : std::get<target_neuron_id>(synapses_[i]);
            if (v == neuron_index) out.push_back(i);
        }
        return out;
    }

    void lock_weights() { is_locked_ = true; }
    void unlock_weights() { is_locked_ = false; }
    bool is_locked() const { return is_locked_; }
    SharedSynapseParameters &get_shared_parameters() { return shared_; }
    const SharedSynapseParameters &get_shared_parameters() const { return shared_; }

private:
    UID uid_;
    UID presynaptic_uid_;
    UID postsynaptic_uid_;
    bool is_locked_ = true;
    std::vector<Synapse> synapses_;
    SharedSynapseParameters shared_;
};

}  // namespace knp::core
