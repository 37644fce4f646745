// knp/core/population.h - a population of neurons of one type.
// Interface of knp/core-library/include/knp/core/population.h:46-285 (constructors from a generator,
// uid, size, indexing, iteration).
#pragma once
#include <functional>
#include <optional>
#include <vector>

#include "uid.h"
#include "../neuron-traits/blifat.h"
#include "../neuron-traits/stdp_synaptic_resource_rule.h"
#include "../neuron-traits/lif.h"
#include "../neuron-traits/altai_lif.h"

namespace knp::core {

template <typename NeuronType>
class Population final {
public:
    using PopulationNeuronType = NeuronType;
    using NeuronParameters = neuron_traits::neuron_parameters<NeuronType>;
    using NeuronGenerator = std::function<std::optional<NeuronParameters>(size_t index)>;
    using NeuronsContainer = std::vector<NeuronParameters>;

    Population(NeuronGenerator generator, size_t neurons_count) : Population(UID(), std::move(generator), neurons_count) {}
    Population(const UID &uid, NeuronGenerator generator, size_t neurons_count) : uid_(uid) {
        neurons_.reserve(neurons_count);
        for (size_t i = 0; i < neurons_count; ++i) {
            auto n = generator(i);
            if (n) neurons_.push_back(*n);
        }
    }

    const UID &get_uid() const { return uid_; }
    size_t size() const { return neurons_.size(); }
    NeuronParameters &operator[](size_t i) { return neurons_[i]; }
    const NeuronParameters &operator[](size_t i) const { return neurons_[i]; }
    auto begin() { return neurons_.begin(); }
    auto end() { return neurons_.end(); }
    auto begin() const { return neurons_.cbegin(); }
    auto end() const { return neurons_.cend(); }
    const NeuronsContainer &get_neurons_parameters() const { return neurons_; }

private:
    UID uid_;
    NeuronsContainer neurons_;
};

}  // namespace knp::core
