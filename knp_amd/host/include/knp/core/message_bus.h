// knp/core/message_bus.h - in-process message bus with per-endpoint typed subscriptions.
// Behaviour of the reference CPU bus (knp/core-library/impl/message_bus_cpu_impl/message_bus_cpu_impl.cpp:
// 41-88, impl/message_endpoint.cpp:77-241): endpoints queue sent messages; route_messages() collects the queues
// in endpoint-creation order and delivers each message to every endpoint subscribed to its sender, keeping
// per-endpoint FIFO order; receive_all_messages() files received messages into the subscriptions
// (type, receiver) whose sender set contains the sender; unload_messages() moves them out.
#pragma once
#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <unordered_set>
#include <utility>
#include <vector>

#include "messaging/messaging.h"

namespace knp::core {

template <class MessageType>
class Subscription {
public:
    using MessageContainerType = std::vector<MessageType>;
    Subscription(const UID &receiver, const std::vector<UID> &senders) : receiver_(receiver) { add_senders(senders); }
    const UID &get_receiver_uid() const { return receiver_; }
    void add_senders(const std::vector<UID> &senders) {
        for (const auto &s : senders) senders_.insert(s);
    }
    bool has_sender(const UID &s) const { return senders_.count(s) != 0; }
    const std::unordered_set<UID, uid_hash> &get_senders() const { return senders_; }
    void add_message(const MessageType &m) { messages_.push_back(m); }
    MessageContainerType &get_messages() { return messages_; }
    void clear_messages() { messages_.clear(); }

private:
    UID receiver_;
    std::unordered_set<UID, uid_hash> senders_;
    MessageContainerType messages_;
};

namespace detail {
struct EndpointState {
    std::mutex mutex;
    std::vector<messaging::MessageVariant> outbox;
    std::vector<messaging::MessageVariant> inbox;  // FIFO
    std::unordered_set<UID, uid_hash> senders;     // union of all subscriptions' senders
};
}  // namespace detail

class MessageBus;

class MessageEndpoint {
public:
    template <class M>
    static constexpr size_t type_index() {
        return std::is_same_v<M, messaging::SpikeMessage> ? 0 : 1;
    }

    template <class M>
    Subscription<M> &subscribe(const UID &receiver, const std::vector<UID> &senders) {
        {
            std::lock_guard<std::mutex> lock(state_->mutex);
            for (const auto &s : senders) state_->senders.insert(s);
        }
        auto key = std::make_pair(type_index<M>(), receiver);
        auto &subs = subs_of<M>();
        auto it = subs.find(key);
        if (it != subs.end()) {
            it->second.add_senders(senders);
            return it->second;
        }
        return subs.emplace(key, Subscription<M>{receiver, senders}).first->second;
    }

    void send_message(const messaging::MessageVariant &m) {
        std::lock_guard<std::mutex> lock(state_->mutex);
        state_->outbox.push_back(m);
    }

    // Files every received message into the matching subscriptions; returns the number of messages.
    size_t receive_all_messages() {
        std::vector<messaging::MessageVariant> got;
        {
            std::lock_guard<std::mutex> lock(state_->mutex);
            got.swap(state_->inbox);
        }
        for (const auto &m : got) {
            const UID &sender = messaging::get_header(m).sender_uid_;
            if (std::holds_alternative<messaging::SpikeMessage>(m)) {
                for (auto &[k, sub] : spike_subs_)
                    if (sub.has_sender(sender)) sub.add_message(std::get<messaging::SpikeMessage>(m));
            } else {
                for (auto &[k, sub] : impact_subs_)
                    if (sub.has_sender(sender)) sub.add_message(std::get<messaging::SynapticImpactMessage>(m));
            }
        }
        return got.size();
    }

    template <class M>
    std::vector<M> unload_messages(const UID &receiver) {
        auto &subs = subs_of<M>();
        auto it = subs.find(std::make_pair(type_index<M>(), receiver));
        if (it == subs.end()) return {};
        std::vector<M> out = std::move(it->second.get_messages());
        it->second.clear_messages();
        return out;
    }

private:
    friend class MessageBus;
    explicit MessageEndpoint(std::shared_ptr<detail::EndpointState> s) : state_(std::move(s)) {}
    template <class M>
    std::map<std::pair<size_t, UID>, Subscription<M>> &subs_of() {
        if constexpr (std::is_same_v<M, messaging::SpikeMessage>) return spike_subs_;
        else return impact_subs_;
    }
    std::shared_ptr<detail::EndpointState> state_;
    std::map<std::pair<size_t, UID>, Subscription<messaging::SpikeMessage>> spike_subs_;
    std::map<std::pair<size_t, UID>, Subscription<messaging::SynapticImpactMessage>> impact_subs_;
};

class MessageBus {
public:
    static MessageBus construct_bus() { return MessageBus(); }
    static MessageBus construct_cpu_bus() { return MessageBus(); }

    MessageEndpoint create_endpoint() {
        auto s = std::make_shared<detail::EndpointState>();
        std::lock_guard<std::mutex> lock(*mutex_);
        endpoints_->push_back(s);
        return MessageEndpoint(s);
    }

    // Delivers every queued message; returns the number of messages routed.
    size_t route_messages() {
        std::lock_guard<std::mutex> lock(*mutex_);
        std::vector<messaging::MessageVariant> all;
        auto &eps = *endpoints_;
        eps.erase(std::remove_if(eps.begin(), eps.end(), [](auto &w) { return w.expired(); }), eps.end());
        for (auto &w : eps) {
            auto s = w.lock();
            std::lock_guard<std::mutex> l2(s->mutex);
            for (auto &m : s->outbox) all.push_back(std::move(m));
            s->outbox.clear();
        }
        for (const auto &m : all) {
            const UID &sender = messaging::get_header(m).sender_uid_;
            for (auto &w : eps) {
                auto s = w.lock();
                std::lock_guard<std::mutex> l2(s->mutex);
                if (s->senders.count(sender)) s->inbox.push_back(m);
            }
        }
        return all.size();
    }

    // True if any endpoint other than `except` subscribes to messages from `sender`.
    bool has_subscriber(const UID &sender, const MessageEndpoint *except = nullptr) const {
        std::lock_guard<std::mutex> lock(*mutex_);
        for (auto &w : *endpoints_) {
            auto s = w.lock();
            if (!s || (except && s == except->state_)) continue;
            std::lock_guard<std::mutex> l2(s->mutex);
            if (s->senders.count(sender)) return true;
        }
        return false;
    }

private:
    std::shared_ptr<std::mutex> mutex_ = std::make_shared<std::mutex>();
    std::shared_ptr<std::vector<std::weak_ptr<detail::EndpointState>>> endpoints_ =
        std::make_shared<std::vector<std::weak_ptr<detail::EndpointState>>>();
};

}  // namespace knp::core
