// mi355x_backend.cpp - knp::core::Backend implementation over the C ABI of libknp_mi355x.so.
#include "knp/backends/gpu-mi355x/backend.h"

#include <cstdio>
#include <stdexcept>
#include <unordered_map>

#include "knp_mi355x.h"

namespace knp::backends::gpu_mi355x {

namespace {

using BLIFAT = neuron_traits::neuron_parameters<neuron_traits::BLIFATNeuron>;
using ResBLIFAT = neuron_traits::neuron_parameters<neuron_traits::SynapticResourceSTDPBLIFATNeuron>;
using LIF = neuron_traits::neuron_parameters<neuron_traits::LIFNeuron>;
using AltAI = neuron_traits::neuron_parameters<neuron_traits::AltAILIF>;

knp_altai_lif_neuron to_abi(const AltAI &n) {
    knp_altai_lif_neuron o;
    o.is_diff = n.is_diff_;
    o.is_reset = n.is_reset_;
    o.leak_rev = n.leak_rev_;
    o.saturate = n.saturate_;
    o.do_not_save = n.do_not_save_;
    o.potential = n.potential_;
    o.pre_impact_potential = n.pre_impact_potential_;
    o.activation_threshold = n.activation_threshold_;
    o.negative_activation_threshold = n.negative_activation_threshold_;
    o.potential_leak = n.potential_leak_;
    o.potential_reset_value = n.potential_reset_value_;
    o.dopamine_value = n.dopamine_value_;
    o.additional_threshold = n.additional_threshold_;
    o.activity_time = n.activity_time_;
    return o;
}

knp_lif_neuron to_abi(const LIF &n) {
    knp_lif_neuron o;
    o.potential = n.potential_;
    o.potential_reset_value = n.potential_reset_value_;
    o.activation_threshold = n.activation_threshold_;
    o.leak_coefficient = n.leak_coefficient_;
    o.refract_counter = n.refract_counter_;
    o.refract_period = n.refract_period_;
    return o;
}

knp_blifat_neuron to_abi(const BLIFAT &n) {
    knp_blifat_neuron o;
    o.n_time_steps_since_last_firing = n.n_time_steps_since_last_firing_;
    o.activation_threshold = n.activation_threshold_;
    o.additional_threshold = n.additional_threshold_;
    o.dynamic_threshold = n.dynamic_threshold_;
    o.threshold_decay = n.threshold_decay_;
    o.threshold_increment = n.threshold_increment_;
    o.postsynaptic_trace = n.postsynaptic_trace_;
    o.postsynaptic_trace_decay = n.postsynaptic_trace_decay_;
    o.postsynaptic_trace_increment = n.postsynaptic_trace_increment_;
    o.inhibitory_conductance = n.inhibitory_conductance_;
    o.inhibitory_conductance_decay = n.inhibitory_conductance_decay_;
    o.potential = n.potential_;
    o.pre_impact_potential = n.pre_impact_potential_;
    o.potential_decay = n.potential_decay_;
    o.bursting_phase = n.bursting_phase_;
    o.bursting_period = n.bursting_period_;
    o.reflexive_weight = n.reflexive_weight_;
    o.reversal_inhibitory_potential = n.reversal_inhibitory_potential_;
    o.absolute_refractory_period = n.absolute_refractory_period_;
    o.potential_reset_value = n.potential_reset_value_;
    o.min_potential = n.min_potential_;
    o.total_blocking_period = n.total_blocking_period_;
    o.dopamine_value = n.dopamine_value_;
    o.stochastic_stimulation = n.stochastic_stimulation_;
    o.random_number_generator_state = n.random_number_generator_state_;
    return o;
}

void from_abi(const knp_blifat_neuron &o, BLIFAT &n) {
    n.n_time_steps_since_last_firing_ = o.n_time_steps_since_last_firing;
    n.additional_threshold_ = o.additional_threshold;
    n.dynamic_threshold_ = o.dynamic_threshold;
    n.postsynaptic_trace_ = o.postsynaptic_trace;
    n.inhibitory_conductance_ = o.inhibitory_conductance;
    n.potential_ = o.potential;
    n.pre_impact_potential_ = o.pre_impact_potential;
    n.bursting_phase_ = o.bursting_phase;
    n.total_blocking_period_ = o.total_blocking_period;
    n.dopamine_value_ = o.dopamine_value;
}

// The resource fields of a synaptic-resource neuron (any base model) into its C ABI struct.
template <class Res, class Out>
void resource_to_abi(const Res &n, Out &o) {
    o.synapse_sum_threshold_coefficient = n.synapse_sum_threshold_coefficient_;
    o.dopamine_plasticity_time = n.dopamine_plasticity_time_;
    o.free_synaptic_resource = n.free_synaptic_resource_;
    o.synaptic_resource_threshold = n.synaptic_resource_threshold_;
    o.resource_drain_coefficient = n.resource_drain_coefficient_;
    o.stability = n.stability_;
    o.stability_change_parameter = n.stability_change_parameter_;
    o.stability_change_at_isi = n.stability_change_at_isi_;
    o.isi_max = n.isi_max_;
    o.d_h = n.d_h_;
    o.isi_status = static_cast<int32_t>(n.isi_status_);
    o.last_step = n.last_step_;
    o.last_spike_step = n.last_spike_step_;
    o.first_isi_spike = n.first_isi_spike_;
    o.is_being_forced = n.is_being_forced_ ? 1 : 0;
}

// The resource state that training changes, from the C ABI struct.
template <class In, class Res>
void resource_from_abi(const In &a, Res &n) {
    n.free_synaptic_resource_ = a.free_synaptic_resource;
    n.stability_ = a.stability;
    n.isi_status_ = static_cast<neuron_traits::ISIPeriodType>(a.isi_status);
    n.last_step_ = a.last_step;
    n.last_spike_step_ = a.last_spike_step;
    n.first_isi_spike_ = a.first_isi_spike;
    n.is_being_forced_ = a.is_being_forced != 0;
}

knp_resource_blifat_neuron to_abi(const ResBLIFAT &n) {
    knp_resource_blifat_neuron o;
    o.base = to_abi(static_cast<const BLIFAT &>(n));
    resource_to_abi(n, o);
    return o;
}

template <class S>
struct SynArrays {
    std::vector<uint32_t> pre, post, delay;
    std::vector<float> w;
    std::vector<uint8_t> type;
    uint32_t max_pre = 0;
    explicit SynArrays(const core::Projection<S> &p) {
        for (const auto &syn : p) {
            const auto &par = std::get<core::synapse_data>(syn);
            pre.push_back((uint32_t)std::get<core::source_neuron_id>(syn));
            post.push_back((uint32_t)std::get<core::target_neuron_id>(syn));
            w.push_back(par.weight_);
            delay.push_back(par.delay_);
            type.push_back((uint8_t)par.output_type_);
            max_pre = std::max(max_pre, pre.back() + 1);
        }
    }
};

}  // namespace

void MI355XBackend::check(int rc, const char *what) const {
    if (rc != KNP_OK) throw std::runtime_error(std::string(what) + ": " + knp_mi355x_last_error());
}

std::shared_ptr<MI355XBackend> MI355XBackend::create() { return std::make_shared<MI355XBackend>(); }

MI355XBackend::MI355XBackend(int device, Semantics semantics) : device_(device), semantics_(semantics) {}

MI355XBackend::~MI355XBackend() {
    if (net_) knp_mi355x_destroy(net_);
}

// The multi-threaded backend runs BLIFATNeuron, SynapticResourceSTDPBLIFATNeuron and LIFNeuron with
// DeltaSynapse and SynapticResourceSTDPDeltaSynapse (cpu-multi-threaded-backend/include/knp/backends/
// cpu-multi-threaded/backend.h:80-88): indexes 0, 1, 4 and 0, 2 of the variants.
std::vector<std::string> MI355XBackend::get_supported_neurons() const {
    if (semantics_ == Semantics::multi_threaded) return {"BLIFATNeuron", "SynapticResourceSTDPBLIFATNeuron", "LIFNeuron"};
    return {"BLIFATNeuron", "SynapticResourceSTDPBLIFATNeuron", "AltAILIF", "SynapticResourceSTDPAltAILIFNeuron",
            "LIFNeuron"};
}

std::vector<std::string> MI355XBackend::get_supported_synapses() const {
    if (semantics_ == Semantics::multi_threaded) return {"DeltaSynapse", "SynapticResourceSTDPDeltaSynapse"};
    return {"DeltaSynapse", "AdditiveSTDPDeltaSynapse", "SynapticResourceSTDPDeltaSynapse"};
}

std::vector<size_t> MI355XBackend::get_supported_population_indexes() const {
    if (semantics_ == Semantics::multi_threaded) return {0, 1, 4};
    std::vector<size_t> out(std::variant_size_v<core::AllPopulationsVariant>);
    for (size_t i = 0; i < out.size(); ++i) out[i] = i;
    return out;
}

std::vector<size_t> MI355XBackend::get_supported_projection_indexes() const {
    if (semantics_ == Semantics::multi_threaded) return {0, 2};
    std::vector<size_t> out(std::variant_size_v<core::AllProjectionsVariant>);
    for (size_t i = 0; i < out.size(); ++i) out[i] = i;
    return out;
}

namespace {
// Value iterators over a snapshot of the network taken by get_network_data (shared by both ends of a range).
template <class Variant>
class SnapshotIterator : public core::Backend::BaseValueIterator<Variant> {
public:
    SnapshotIterator(std::shared_ptr<const std::vector<Variant>> data, size_t pos) : data_(std::move(data)), pos_(pos) {}
    Variant operator*() const override { return (*data_)[pos_]; }
    core::Backend::BaseValueIterator<Variant> &operator++() override {
        ++pos_;
        return *this;
    }
    bool operator==(const core::Backend::BaseValueIterator<Variant> &rhs) const override {
        if (typeid(*this) != typeid(rhs)) return false;
        const auto &r = dynamic_cast<const SnapshotIterator &>(rhs);
        return r.data_ == data_ && r.pos_ == pos_;
    }

private:
    std::shared_ptr<const std::vector<Variant>> data_;
    size_t pos_;
};
}  // namespace

core::Backend::DataRanges MI355XBackend::get_network_data() const {
    auto copy = network_copy();
    auto pops = std::make_shared<const std::vector<PopulationVariants>>(std::move(copy.first));
    auto projs = std::make_shared<const std::vector<ProjectionVariants>>(std::move(copy.second));
    using PopIt = SnapshotIterator<core::AllPopulationsVariant>;
    using ProjIt = SnapshotIterator<core::AllProjectionsVariant>;
    DataRanges r;
    r.population_range = {std::make_unique<PopIt>(pops, 0), std::make_unique<PopIt>(pops, pops->size())};
    r.projection_range = {std::make_unique<ProjIt>(projs, 0), std::make_unique<ProjIt>(projs, projs->size())};
    return r;
}

void MI355XBackend::use_device(const core::UID &uid) {
    int n = 0;
    check(knp_mi355x_device_count(&n), "device count");
    int index = -1;
    auto devs = get_devices();
    for (int d = 0; d < (int)devs.size(); ++d)
        if (devs[d]->get_uid() == uid) index = d;
    if (index < 0) throw std::logic_error("Not all devices were selected.");
    if (index == device_) return;
    if (net_) throw std::logic_error("The device must be selected before the backend is initialized.");
    device_ = index;
}

void MI355XBackend::select_devices(const std::set<core::UID> &uids) {
    core::Backend::select_devices(uids);
    if (!get_current_devices().empty()) use_device(get_current_devices().back()->get_uid());
}

void MI355XBackend::select_device(std::unique_ptr<core::Device> &&device) {
    if (!device) throw std::logic_error("Not all devices were selected.");
    use_device(device->get_uid());
    core::Backend::select_device(std::move(device));
}

void MI355XBackend::load_populations(const std::vector<PopulationVariants> &populations) {
    if (net_) throw std::logic_error("Populations must be loaded before the backend is initialized.");
    populations_ = populations;
}

void MI355XBackend::load_projections(const std::vector<ProjectionVariants> &projections) {
    if (net_) throw std::logic_error("Projections must be loaded before the backend is initialized.");
    projections_ = projections;
}

void MI355XBackend::remove_projections(const std::vector<core::UID> &uids) {
    if (net_) throw std::logic_error("Projections cannot be removed after initialization.");
    for (const auto &u : uids)
        projections_.erase(std::remove_if(projections_.begin(), projections_.end(),
                                          [&](const auto &v) { return std::visit([](auto &p) { return p.get_uid(); }, v) == u; }),
                           projections_.end());
}

void MI355XBackend::remove_populations(const std::vector<core::UID> &uids) {
    if (net_) throw std::logic_error("Populations cannot be removed after initialization.");
    for (const auto &u : uids)
        populations_.erase(std::remove_if(populations_.begin(), populations_.end(),
                                          [&](const auto &v) { return std::visit([](auto &p) { return p.get_uid(); }, v) == u; }),
                           populations_.end());
}

std::vector<std::unique_ptr<core::Device>> MI355XBackend::get_devices() const {
    std::vector<std::unique_ptr<core::Device>> out;
    int n = 0;
    if (knp_mi355x_device_count(&n) != KNP_OK) return out;
    for (int d = 0; d < n; ++d) {
        char name[256];
        if (knp_mi355x_device_name(d, name, sizeof name) != KNP_OK) continue;
        core::UID uid(false);
        auto &b = const_cast<std::array<uint8_t, 16> &>(uid.bytes());
        b[0] = 0x4D;  // deterministic per device index
        b[15] = (uint8_t)(d + 1);
        out.push_back(std::make_unique<core::Device>(name, uid));
    }
    return out;
}

void MI355XBackend::start_learning() {
    learning_ = true;
    for (auto &v : projections_) std::visit([](auto &p) { p.unlock_weights(); }, v);
    if (net_) check(knp_mi355x_start_learning(net_), "start_learning");
}

void MI355XBackend::stop_learning() {
    learning_ = false;
    for (auto &v : projections_) std::visit([](auto &p) { p.lock_weights(); }, v);
    if (net_) check(knp_mi355x_stop_learning(net_), "stop_learning");
}

void MI355XBackend::_init() {
    if (net_) return;
    check(knp_mi355x_create(device_, &net_), "create");
    if (semantics_ == Semantics::multi_threaded)
        check(knp_mi355x_set_message_semantics(net_, KNP_SEMANTICS_MULTI_THREADED), "message semantics");
    std::unordered_map<core::UID, uint32_t, core::uid_hash> pop_index;
    for (const auto &v : populations_) {
        uint32_t idx = 0;
        if (auto *p = std::get_if<core::Population<neuron_traits::BLIFATNeuron>>(&v)) {
            std::vector<knp_blifat_neuron> a;
            for (const auto &n : *p) a.push_back(to_abi(n));
            check(knp_mi355x_add_blifat_population(net_, a.data(), (uint32_t)a.size(), &idx), "load population");
            pop_uids_.push_back(p->get_uid());
        } else if (auto *pl = std::get_if<core::Population<neuron_traits::LIFNeuron>>(&v)) {
            std::vector<knp_lif_neuron> a;
            for (const auto &n : *pl) a.push_back(to_abi(n));
            check(knp_mi355x_add_lif_population(net_, a.data(), (uint32_t)a.size(), &idx), "load population");
            pop_uids_.push_back(pl->get_uid());
        } else if (auto *pa = std::get_if<core::Population<neuron_traits::AltAILIF>>(&v)) {
            std::vector<knp_altai_lif_neuron> a;
            for (const auto &n : *pa) a.push_back(to_abi(n));
            check(knp_mi355x_add_altai_lif_population(net_, a.data(), (uint32_t)a.size(), &idx), "load population");
            pop_uids_.push_back(pa->get_uid());
        } else if (auto *pr = std::get_if<core::Population<neuron_traits::SynapticResourceSTDPAltAILIFNeuron>>(&v)) {
            std::vector<knp_resource_altai_lif_neuron> a;
            for (const auto &n : *pr) {
                knp_resource_altai_lif_neuron o;
                o.base = to_abi(static_cast<const AltAI &>(n));
                resource_to_abi(n, o);
                a.push_back(o);
            }
            check(knp_mi355x_add_resource_altai_lif_population(net_, a.data(), (uint32_t)a.size(), &idx),
                  "load population");
            pop_uids_.push_back(pr->get_uid());
        } else {
            const auto &p2 = std::get<core::Population<neuron_traits::SynapticResourceSTDPBLIFATNeuron>>(v);
            std::vector<knp_resource_blifat_neuron> a;
            for (const auto &n : p2) a.push_back(to_abi(n));
            check(knp_mi355x_add_resource_blifat_population(net_, a.data(), (uint32_t)a.size(), &idx),
                  "load population");
            pop_uids_.push_back(p2.get_uid());
        }
        pop_index[pop_uids_.back()] = idx;
    }
    auto &ep = get_message_endpoint();
    for (const auto &v : projections_) {
        std::visit(
            [&](const auto &p) {
                using P = std::decay_t<decltype(p)>;
                using S = typename P::ProjectionSynapseType;
                SynArrays<S> a(p);
                auto it_post = pop_index.find(p.get_postsynaptic());
                if (it_post == pop_index.end())
                    throw std::runtime_error("Projection targets a population that is not loaded.");
                auto it_pre = pop_index.find(p.get_presynaptic());
                const uint32_t pre_pop = it_pre == pop_index.end() ? KNP_NO_POPULATION : it_pre->second;
                uint32_t q = 0;
                const uint64_t n = a.pre.size();
                if constexpr (std::is_same_v<S, synapse_traits::DeltaSynapse>) {
                    check(knp_mi355x_add_delta_projection(net_, pre_pop, a.max_pre, it_post->second, a.pre.data(),
                                                          a.post.data(), a.w.data(), a.delay.data(), a.type.data(), n,
                                                          &q),
                          "load projection");
                } else if constexpr (std::is_same_v<S, synapse_traits::SynapticResourceSTDPDeltaSynapse>) {
                    std::vector<knp_resource_rule> rules;
                    for (const auto &syn : p) {
                        const auto &r = std::get<core::synapse_data>(syn).rule_;
                        knp_resource_rule o;
                        o.synaptic_resource = r.synaptic_resource_;
                        o.w_min = r.w_min_;
                        o.w_max = r.w_max_;
                        o.d_u = r.d_u_;
                        o.dopamine_plasticity_period = r.dopamine_plasticity_period_;
                        o.last_spike_step = r.last_spike_step_;
                        o.had_hebbian_update = r.had_hebbian_update_;
                        o.has_contributed = r.has_contributed_;
                        rules.push_back(o);
                    }
                    check(knp_mi355x_add_resource_stdp_projection(net_, pre_pop, a.max_pre, it_post->second,
                                                                  a.pre.data(), a.post.data(), a.w.data(),
                                                                  a.delay.data(), a.type.data(), rules.data(), n, &q),
                          "load projection");
                } else {
                    std::vector<float> tp, tm;
                    for (const auto &syn : p) {
                        tp.push_back(std::get<core::synapse_data>(syn).rule_.tau_plus_);
                        tm.push_back(std::get<core::synapse_data>(syn).rule_.tau_minus_);
                    }
                    std::vector<uint32_t> sp;
                    std::vector<uint8_t> sm;
                    for (const auto &[uid, mode] : p.get_shared_parameters().stdp_populations_) {
                        auto it = pop_index.find(uid);
                        if (it == pop_index.end()) continue;
                        sp.push_back(it->second);
                        sm.push_back(mode == P::SharedSynapseParameters::ProcessingType::STDPOnly ? KNP_STDP_ONLY
                                                                                                   : KNP_STDP_AND_SPIKE);
                    }
                    check(knp_mi355x_add_additive_stdp_projection(
                              net_, pre_pop, a.max_pre, it_post->second, a.pre.data(), a.post.data(), a.w.data(),
                              a.delay.data(), a.type.data(), tp.data(), tm.data(), n, sp.data(), sm.data(),
                              (uint32_t)sp.size(), &q),
                          "load projection");
                }
                check(knp_mi355x_set_locked(net_, q, p.is_locked() ? 1 : 0), "lock state");
                proj_uids_.push_back(p.get_uid());
                // presynaptic senders that are not populations of this backend reach the projection over the
                // bus (cpu-library init.h:104-109)
                if (p.get_presynaptic() && pre_pop == KNP_NO_POPULATION)
                    ep.subscribe<core::messaging::SpikeMessage>(p.get_uid(), {p.get_presynaptic()});
            },
            v);
    }
    check(knp_mi355x_init(net_), "init");
}

void MI355XBackend::_step() {
    if (!net_) _init();
    auto &bus = get_message_bus();
    auto &ep = get_message_endpoint();
    bus.route_messages();
    ep.receive_all_messages();
    // channel messages for projections -> device
    for (uint32_t q = 0; q < proj_uids_.size(); ++q) {
        for (const auto &m : ep.unload_messages<core::messaging::SpikeMessage>(proj_uids_[q])) {
            check(knp_mi355x_push_spikes(net_, q, m.neuron_indexes_.data(), (uint32_t)m.neuron_indexes_.size()),
                  "spike message");
        }
    }
    check(knp_mi355x_step(net_, 1), "step");
    // population spikes of this step for observers (populations_old.h:110-118: only when non-empty)
    for (uint32_t p = 0; p < pop_uids_.size(); ++p) {
        if (!bus.has_subscriber(pop_uids_[p], &ep)) continue;
        const size_t n = std::visit([](const auto &x) { return x.size(); }, populations_[p]);
        core::messaging::SpikeMessage msg{{pop_uids_[p], get_step()}, {}};
        msg.neuron_indexes_.resize(n);
        uint32_t cnt = 0;
        check(knp_mi355x_read_last_spikes(net_, p, msg.neuron_indexes_.data(), (uint32_t)n, &cnt), "read spikes");
        msg.neuron_indexes_.resize(cnt);
        if (cnt) ep.send_message(msg);
    }
    bus.route_messages();
    ep.receive_all_messages();
    gad_step();
}

std::pair<std::vector<MI355XBackend::PopulationVariants>, std::vector<MI355XBackend::ProjectionVariants>>
MI355XBackend::network_copy() const {
    auto pops = populations_;
    auto projs = projections_;
    if (!net_) return {pops, projs};
    for (uint32_t i = 0; i < pops.size(); ++i) {
        if (auto *p = std::get_if<core::Population<neuron_traits::BLIFATNeuron>>(&pops[i])) {
            std::vector<knp_blifat_neuron> a(p->size());
            check(knp_mi355x_read_blifat_neurons(net_, i, a.data(), (uint32_t)a.size()), "read neurons");
            for (size_t k = 0; k < a.size(); ++k) from_abi(a[k], (*p)[k]);
        } else if (auto *pl = std::get_if<core::Population<neuron_traits::LIFNeuron>>(&pops[i])) {
            std::vector<knp_lif_neuron> a(pl->size());
            check(knp_mi355x_read_lif_neurons(net_, i, a.data(), (uint32_t)a.size()), "read neurons");
            for (size_t k = 0; k < a.size(); ++k) {
                (*pl)[k].potential_ = a[k].potential;
                (*pl)[k].refract_counter_ = a[k].refract_counter;
            }
        } else if (auto *pa = std::get_if<core::Population<neuron_traits::AltAILIF>>(&pops[i])) {
            std::vector<knp_altai_lif_neuron> a(pa->size());
            check(knp_mi355x_read_altai_lif_neurons(net_, i, a.data(), (uint32_t)a.size()), "read neurons");
            for (size_t k = 0; k < a.size(); ++k) {
                (*pa)[k].potential_ = a[k].potential;
                (*pa)[k].pre_impact_potential_ = a[k].pre_impact_potential;
                (*pa)[k].dopamine_value_ = a[k].dopamine_value;
                (*pa)[k].activity_time_ = a[k].activity_time;
            }
        } else if (auto *pr = std::get_if<core::Population<neuron_traits::SynapticResourceSTDPAltAILIFNeuron>>(&pops[i])) {
            std::vector<knp_resource_altai_lif_neuron> a(pr->size());
            check(knp_mi355x_read_resource_altai_lif_neurons(net_, i, a.data(), (uint32_t)a.size()), "read neurons");
            for (size_t k = 0; k < a.size(); ++k) {
                auto &n = (*pr)[k];
                n.potential_ = a[k].base.potential;
                n.pre_impact_potential_ = a[k].base.pre_impact_potential;
                n.dopamine_value_ = a[k].base.dopamine_value;
                n.additional_threshold_ = a[k].base.additional_threshold;
                n.activity_time_ = a[k].base.activity_time;
                resource_from_abi(a[k], n);
            }
        } else {
            auto &p2 = std::get<core::Population<neuron_traits::SynapticResourceSTDPBLIFATNeuron>>(pops[i]);
            std::vector<knp_resource_blifat_neuron> a(p2.size());
            check(knp_mi355x_read_resource_neurons(net_, i, a.data(), (uint32_t)a.size()), "read neurons");
            for (size_t k = 0; k < a.size(); ++k) {
                from_abi(a[k].base, p2[k]);
                resource_from_abi(a[k], p2[k]);
            }
        }
    }
    for (uint32_t q = 0; q < projs.size(); ++q) {
        std::visit(
            [&](auto &p) {
                using S = typename std::decay_t<decltype(p)>::ProjectionSynapseType;
                std::vector<float> w(p.size());
                check(knp_mi355x_read_weights(net_, q, w.data(), w.size()), "read weights");
                for (size_t s = 0; s < w.size(); ++s) std::get<core::synapse_data>(p[s]).weight_ = w[s];
                if constexpr (std::is_same_v<S, synapse_traits::SynapticResourceSTDPDeltaSynapse>) {
                    std::vector<knp_resource_rule> r(p.size());
                    check(knp_mi355x_read_resource_rules(net_, q, r.data(), r.size()), "read rules");
                    for (size_t s = 0; s < r.size(); ++s) {
                        auto &rule = std::get<core::synapse_data>(p[s]).rule_;
                        rule.synaptic_resource_ = r[s].synaptic_resource;
                        rule.last_spike_step_ = r[s].last_spike_step;
                        rule.had_hebbian_update_ = r[s].had_hebbian_update != 0;
                        rule.has_contributed_ = r[s].has_contributed != 0;
                    }
                }
            },
            projs[q]);
    }
    return {pops, projs};
}

}  // namespace knp::backends::gpu_mi355x
