// knp/core/backend.h - the backend interface.
// Interface of knp/core-library/include/knp/core/backend.h:55-457 for the populations and projections this
// build supports: the same pure virtuals (plasticity_supported :103, get_supported_neurons :109,
// get_supported_synapses :115, get_supported_population_indexes :121, get_supported_projection_indexes :127,
// load_all_projections :137, load_all_populations :146, remove_* :153,160, get_devices :172,
// stop_learning :295, start_learning :300, _init :314, _step :321, get_network_data :420 with its
// BaseValueIterator / DataRanges types :336-413), the device selection (select_devices :200, select_device
// :207, get_current_devices :183,193; knp/core-library/impl/backend.cpp:166-192), the same run loop
// (start/stop/get_step, knp/core-library/impl/backend.cpp:57-141: _init once, then _step while running,
// stop() from any thread, exceptions reset the running flag and propagate) and bus access (subscribe :222).
#pragma once
#include <atomic>
#include <functional>
#include <iterator>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <variant>
#include <vector>

#include "message_bus.h"
#include "population.h"
#include "projection.h"

namespace knp::core {

using AllPopulationsVariant = std::variant<Population<neuron_traits::BLIFATNeuron>,
                                           Population<neuron_traits::SynapticResourceSTDPBLIFATNeuron>,
                                           Population<neuron_traits::AltAILIF>,
                                           Population<neuron_traits::SynapticResourceSTDPAltAILIFNeuron>,
                                           Population<neuron_traits::LIFNeuron>>;
using AllProjectionsVariant = std::variant<Projection<synapse_traits::DeltaSynapse>,
                                           Projection<synapse_traits::AdditiveSTDPDeltaSynapse>,
                                           Projection<synapse_traits::SynapticResourceSTDPDeltaSynapse>>;

// A compute device (knp/core-library/include/knp/core/device.h:37-109).
class Device {
public:
    Device(std::string name, UID uid) : name_(std::move(name)), uid_(uid) {}
    virtual ~Device() = default;
    const std::string &get_name() const { return name_; }
    const UID &get_uid() const { return uid_; }

private:
    std::string name_;
    UID uid_;
};

class Backend {
public:
    using RunPredicate = std::function<bool(Step)>;

    virtual ~Backend() = default;

    virtual bool plasticity_supported() const = 0;
    virtual std::vector<std::string> get_supported_neurons() const = 0;
    virtual std::vector<std::string> get_supported_synapses() const = 0;
    // Indexes into AllPopulationsVariant / AllProjectionsVariant of the types the backend runs.
    virtual std::vector<size_t> get_supported_population_indexes() const = 0;
    virtual std::vector<size_t> get_supported_projection_indexes() const = 0;
    virtual void load_all_projections(const std::vector<AllProjectionsVariant> &projections) = 0;
    virtual void load_all_populations(const std::vector<AllPopulationsVariant> &populations) = 0;
    virtual void remove_projections(const std::vector<UID> &uids) = 0;
    virtual void remove_populations(const std::vector<UID> &uids) = 0;
    virtual std::vector<std::unique_ptr<Device>> get_devices() const = 0;
    std::vector<std::unique_ptr<Device>> &get_current_devices() { return devices_; }
    const std::vector<std::unique_ptr<Device>> &get_current_devices() const { return devices_; }
    // As the reference: the first listed device whose UID is in `uids` is selected; the call fails when that
    // does not select as many devices as were asked for.
    virtual void select_devices(const std::set<UID> &uids) {
        for (auto &&device : get_devices()) {
            if (uids.find(device->get_uid()) != uids.end()) {
                devices_.push_back(std::move(device));
                return;
            }
        }
        if (uids.size() != devices_.size()) throw std::logic_error("Not all devices were selected.");
    }
    virtual void select_device(std::unique_ptr<Device> &&device) {
        devices_.clear();
        devices_.push_back(std::move(device));
    }

    // Access to the network by value: iterators over populations or projections that yield copies.
    template <class Type>
    class BaseValueIterator {
    public:
        using iterator_category = std::input_iterator_tag;
        using value_type = Type;
        BaseValueIterator() = default;
        virtual Type operator*() const = 0;
        virtual BaseValueIterator &operator++() = 0;
        virtual bool operator==(const BaseValueIterator &rhs) const = 0;
        virtual bool operator!=(const BaseValueIterator &rhs) const { return !(*this == rhs); }
        virtual ~BaseValueIterator() = default;
    };
    struct DataRanges {
        std::pair<std::unique_ptr<BaseValueIterator<AllProjectionsVariant>>,
                  std::unique_ptr<BaseValueIterator<AllProjectionsVariant>>>
            projection_range;
        std::pair<std::unique_ptr<BaseValueIterator<AllPopulationsVariant>>,
                  std::unique_ptr<BaseValueIterator<AllPopulationsVariant>>>
            population_range;
    };
    // Iterator ranges over copies of the populations and projections with their current state.
    virtual DataRanges get_network_data() const = 0;

    template <typename MessageType>
    Subscription<MessageType> &subscribe(const UID &receiver, const std::vector<UID> &senders) {
        return message_endpoint_.subscribe<MessageType>(receiver, senders);
    }
    MessageBus &get_message_bus() { return *message_bus_; }
    virtual MessageEndpoint &get_message_endpoint() { return message_endpoint_; }

    void start() { start(RunPredicate{}, RunPredicate{}); }
    void start(const RunPredicate &run_predicate) { start(run_predicate, RunPredicate{}); }
    void start(const RunPredicate &pre_step, const RunPredicate &post_step) {
        try {
            pre_start();
            started_ = true;
            while (started_) {
                if (pre_step && !pre_step(get_step())) break;
                _step();
                if (post_step && !post_step(get_step())) break;
            }
            started_ = false;
        } catch (...) {
            started_ = false;
            throw;
        }
    }
    void stop() { started_ = false; }
    Step get_step() const { return step_; }
    bool running() const { return started_; }

    virtual void stop_learning() = 0;
    virtual void start_learning() = 0;
    virtual void _init() = 0;
    virtual void _step() = 0;
    virtual void _uninit() {}

protected:
    Backend() : message_bus_(std::make_shared<MessageBus>()), message_endpoint_(message_bus_->create_endpoint()) {}
    explicit Backend(std::shared_ptr<MessageBus> bus)
        : message_bus_(std::move(bus)), message_endpoint_(message_bus_->create_endpoint()) {}
    void pre_start() {
        if (!initialized_) {
            _init();
            initialized_ = true;
        }
    }
    Step gad_step() { return step_++; }
    bool initialized_ = false;

private:
    std::vector<std::unique_ptr<Device>> devices_;
    std::shared_ptr<MessageBus> message_bus_;
    MessageEndpoint message_endpoint_;
    std::atomic<bool> started_{false};
    Step step_ = 0;
};

}  // namespace knp::core
