// knp/core/messaging/messaging.h - bus messages.
// Interfaces of knp/core-library/include/knp/core/messaging/{message_header.h:39-49, spike_message.h:40-67,
// synaptic_impact_message.h:45-123}: header {sender uid, send time}, SpikeMessage {header, neuron indexes},
// SynapticImpact, SynapticImpactMessage {header, pre/post population uids, is_forcing, impacts}.
#pragma once
#include <cstdint>
#include <variant>
#include <vector>

#include "../uid.h"
#include "../../synapse-traits/output_types.h"

namespace knp::core::messaging {

struct MessageHeader {
    UID sender_uid_;
    Step send_time_ = 0;
};

using SpikeIndex = uint32_t;
using SpikeData = std::vector<SpikeIndex>;

struct SpikeMessage {
    MessageHeader header_;
    SpikeData neuron_indexes_;
};

struct SynapticImpact {
    uint64_t connection_index_;
    float impact_value_;
    knp::synapse_traits::OutputType synapse_type_;
    uint32_t presynaptic_neuron_index_;
    uint32_t postsynaptic_neuron_index_;
};

struct SynapticImpactMessage {
    MessageHeader header_;
    UID presynaptic_population_uid_;
    UID postsynaptic_population_uid_;
    bool is_forcing_ = false;
    std::vector<SynapticImpact> impacts_;
};

using MessageVariant = std::variant<SpikeMessage, SynapticImpactMessage>;

inline const MessageHeader &get_header(const MessageVariant &m) {
    return std::visit([](const auto &x) -> const MessageHeader & { return x.header_; }, m);
}

}  // namespace knp::core::messaging
