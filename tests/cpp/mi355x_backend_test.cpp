// mi355x_backend_test.cpp - the reference's backend integration tests, run against the MI355X backend
// through the knp::core API: SmallestNetwork, AdditiveSTDPBLIFATNetwork, ResourceSTDPBLIFATNetwork and
// ResourceSTDPAltAILIFNetwork (knp/tests/backend/single_thread_cpu_test.cpp:54-358, generators
// knp/tests/common/generators.h:48-100),
// a LIF neuron trace (lif_impl.h:31-84), an AltAI-LIF known answer (knp/tests/framework/altai_lif_test.cpp),
// the supported-type names (:362-381), BackendLoader (knp/tests/framework/backend_loader_test.cpp:29-70)
// and the Backend::start run loop with a stop predicate.
// The multi-threaded cases load libknp-mi355x-mt-backend.so from the same folder and run the reference's
// MultiThreadCpuSuite cases (knp/tests/backend/multi_thread_cpu_test.cpp:54-160) and a repeated-index message,
// which that backend expands as one weight multiple (cpu-library delta_impl.h:144).
// Usage: mi355x_backend_test <path to libknp-mi355x-backend.so>. Prints one PASS/FAIL line per case.
#include <algorithm>
#include <cstdio>
#include <filesystem>
#include <functional>
#include <string>
#include <vector>

#include "knp/backends/gpu-mi355x/backend.h"
#include "knp/framework/backend_loader.h"

using knp::core::UID;
using knp::core::messaging::SpikeMessage;
using DeltaProjection = knp::core::Projection<knp::synapse_traits::DeltaSynapse>;
using BLIFATPopulation = knp::core::Population<knp::neuron_traits::BLIFATNeuron>;
using Backend = knp::backends::gpu_mi355x::MI355XBackend;

static int g_failures = 0;
#define CHECK(cond)                                                              \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::printf("  check failed at line %d: %s\n", __LINE__, #cond);     \
            throw std::runtime_error("check failed");                            \
        }                                                                        \
    } while (0)

static void run_case(const char *name, const std::function<void()> &f) {
    try {
        f();
        std::printf("PASS %s\n", name);
    } catch (const std::exception &e) {
        std::printf("FAIL %s: %s\n", name, e.what());
        ++g_failures;
    }
    std::fflush(stdout);
}

static const std::vector<knp::core::Step> kExpected = {1, 6, 7, 11, 12, 13, 16, 17, 18, 19};

// The network of a backend as get_network_copy builds it (knp/base-framework/impl/synchronization.cpp:26-41):
// walk Backend::get_network_data()'s ranges through a const Backend and take every population and projection.
static std::pair<std::vector<knp::core::AllPopulationsVariant>, std::vector<knp::core::AllProjectionsVariant>>
network_copy(const knp::core::Backend &backend) {
    auto data_ranges = backend.get_network_data();
    std::pair<std::vector<knp::core::AllPopulationsVariant>, std::vector<knp::core::AllProjectionsVariant>> res;
    for (auto &iter = *data_ranges.population_range.first; iter != *data_ranges.population_range.second; ++iter)
        res.first.push_back(*iter);
    for (auto &iter = *data_ranges.projection_range.first; iter != *data_ranges.projection_range.second; ++iter)
        res.second.push_back(*iter);
    return res;
}

// Drive a network for 20 steps: input spikes on steps 0, 5, 10, 15; record steps with output spikes.
static std::vector<knp::core::Step> run_smallest(knp::core::Backend &backend, const UID &input_proj,
                                                 const UID &population) {
    auto endpoint = backend.get_message_bus().create_endpoint();
    const UID in_channel_uid, out_channel_uid;
    backend.subscribe<SpikeMessage>(input_proj, {in_channel_uid});
    endpoint.subscribe<SpikeMessage>(out_channel_uid, {population});
    std::vector<knp::core::Step> results;
    for (knp::core::Step step = 0; step < 20; ++step) {
        if (step % 5 == 0) endpoint.send_message(SpikeMessage{{in_channel_uid, 0}, {0}});
        backend._step();
        endpoint.receive_all_messages();
        auto output = endpoint.unload_messages<SpikeMessage>(out_channel_uid);
        if (!output.empty()) results.push_back(step);
    }
    return results;
}

// ResourceSTDPBLIFATNetwork / ResourceSTDPAltAILIFNetwork (single_thread_cpu_test.cpp:179-358): one
// synaptic-resource neuron of the given model with a plastic input and a plastic delay-6 loop; same spike
// steps as SmallestNetwork, and training changes the loop weight.
template <class NeuronType>
static void run_resource_network() {
    using STDPDeltaProjection = knp::core::Projection<knp::synapse_traits::SynapticResourceSTDPDeltaSynapse>;
    using StdpPopulation = knp::core::Population<NeuronType>;
    Backend backend;
    StdpPopulation population{
        UID(),
        [](uint64_t) -> std::optional<typename StdpPopulation::NeuronParameters> {
            typename StdpPopulation::NeuronParameters neuron{{}};
            neuron.synaptic_resource_threshold_ = 1;
            neuron.free_synaptic_resource_ = 2;
            neuron.isi_max_ = 0;
            return neuron;
        },
        1};
    auto loop = STDPDeltaProjection{
        population.get_uid(), population.get_uid(),
        [](size_t) { return STDPDeltaProjection::Synapse{{{1.0, 6, knp::synapse_traits::OutputType::EXCITATORY}, {0, 1, 2}}, 0, 0}; },
        1};
    STDPDeltaProjection input{
        UID{false}, population.get_uid(),
        [](size_t) { return STDPDeltaProjection::Synapse{{{1.0, 1, knp::synapse_traits::OutputType::EXCITATORY}, {0, 1, 2, 0.1F}}, 0, 0}; },
        1};
    const UID input_uid = input.get_uid();
    backend.load_populations({population});
    backend.load_projections({input, loop});
    backend._init();
    backend.start_learning();
    CHECK(run_smallest(backend, input_uid, population.get_uid()) == kExpected);
    auto data = network_copy(backend);
    const auto &loop_after = std::get<STDPDeltaProjection>(data.second[1]);
    CHECK(std::get<knp::core::synapse_data>(loop_after[0]).weight_ != 1.0f);
}

// One message {0, 1, 0, 0} into two default BLIFAT neurons (weight float(0.1), delay 1, threshold 0.300000008):
// three impacts of float(0.1) sum to 40265319 * 2^-27 < threshold in the single-threaded backend; the
// multi-threaded one sends float(3 * float(0.1)) = 10066330 * 2^-25 >= threshold. Returns the neurons that
// fired at step 1.
static std::vector<uint32_t> repeated_index_spikes(knp::core::Backend &backend) {
    BLIFATPopulation population{[](size_t) {
                                    knp::neuron_traits::neuron_parameters<knp::neuron_traits::BLIFATNeuron> n;
                                    n.activation_threshold_ = 0.3000000080;
                                    return n;
                                },
                                2};
    DeltaProjection input{UID{false}, population.get_uid(),
                          [](size_t i) { return DeltaProjection::Synapse{{0.1f, 1, knp::synapse_traits::OutputType::EXCITATORY}, i, i}; },
                          2};
    const UID input_uid = input.get_uid();
    backend.load_all_populations({population});
    backend.load_all_projections({input});
    auto endpoint = backend.get_message_bus().create_endpoint();
    const UID in_channel_uid, out_channel_uid;
    backend.subscribe<SpikeMessage>(input_uid, {in_channel_uid});
    endpoint.subscribe<SpikeMessage>(out_channel_uid, {population.get_uid()});
    endpoint.send_message(SpikeMessage{{in_channel_uid, 0}, {0, 1, 0, 0}});
    std::vector<uint32_t> fired;
    for (int step = 0; step < 2; ++step) {
        backend._step();
        endpoint.receive_all_messages();
        for (const auto &m : endpoint.unload_messages<SpikeMessage>(out_channel_uid))
            if (step == 1) fired.insert(fired.end(), m.neuron_indexes_.begin(), m.neuron_indexes_.end());
    }
    return fired;
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <libknp-mi355x-backend.so>\n", argv[0]);
        return 2;
    }
    const std::filesystem::path lib = argv[1];
    // --cpu-only: the cases that need no HIP device (type names, loader resolution)
    const bool cpu_only = argc > 2 && std::string(argv[2]) == "--cpu-only";
    const std::filesystem::path mt_lib = lib.parent_path() / "libknp-mi355x-mt-backend.so";
    // NeuronsGettingTest / SynapsesGettingTest (multi_thread_cpu_test.cpp:141-160) through the loader
    run_case("MultiThreadedTypesGetting", [&] {
        knp::framework::BackendLoader loader;
        CHECK(knp::framework::BackendLoader::is_backend(mt_lib));
        auto backend = loader.load(mt_lib);
        const auto neurons = backend->get_supported_neurons();
        const auto synapses = backend->get_supported_synapses();
        CHECK(neurons.size() <= std::variant_size_v<knp::core::AllPopulationsVariant> && neurons[0] == "BLIFATNeuron");
        CHECK(synapses.size() <= std::variant_size_v<knp::core::AllProjectionsVariant> && synapses[0] == "DeltaSynapse");
        CHECK((backend->get_supported_population_indexes() == std::vector<size_t>{0, 1, 4}));
        CHECK((backend->get_supported_projection_indexes() == std::vector<size_t>{0, 2}));
    });
    if (cpu_only) {
        run_case("NeuronsAndSynapsesGetting", [] {
            Backend backend;
            CHECK(backend.get_supported_neurons()[0] == "BLIFATNeuron");
            CHECK(backend.get_supported_synapses()[0] == "DeltaSynapse");
        });
        run_case("BackendLoaderResolves", [&] {
            knp::framework::BackendLoader loader;
            CHECK(knp::framework::BackendLoader::is_backend(lib));
            CHECK(!knp::framework::BackendLoader::is_backend(lib.parent_path() / "no-such-backend"));
            auto backend = loader.load(lib);
            CHECK(backend->get_supported_synapses().size() == 3);
            bool threw = false;
            try {
                loader.load(lib.parent_path() / "no-such-backend");
            } catch (const std::runtime_error &) {
                threw = true;
            }
            CHECK(threw);
        });
        return g_failures ? 1 : 0;
    }

    run_case("SmallestNetwork", [] {
        Backend backend;
        BLIFATPopulation population{[](size_t) { return knp::neuron_traits::neuron_parameters<knp::neuron_traits::BLIFATNeuron>{}; }, 1};
        DeltaProjection loop{population.get_uid(), population.get_uid(),
                             [](size_t) { return DeltaProjection::Synapse{{1.1, 6, knp::synapse_traits::OutputType::EXCITATORY}, 0, 0}; },
                             1};
        DeltaProjection input{UID{false}, population.get_uid(),
                              [](size_t) { return DeltaProjection::Synapse{{1.0, 1, knp::synapse_traits::OutputType::EXCITATORY}, 0, 0}; },
                              1};
        const UID input_uid = input.get_uid();
        backend.load_populations({population});
        backend.load_projections({input, loop});
        backend._init();
        CHECK(run_smallest(backend, input_uid, population.get_uid()) == kExpected);
        CHECK(backend.get_step() == 20);
    });

    run_case("AdditiveSTDPBLIFATNetwork", [] {
        using STDPDeltaProjection = knp::core::Projection<knp::synapse_traits::AdditiveSTDPDeltaSynapse>;
        Backend backend;
        BLIFATPopulation population{UID(), [](size_t) { return knp::neuron_traits::neuron_parameters<knp::neuron_traits::BLIFATNeuron>{}; }, 1};
        auto loop = STDPDeltaProjection{
            population.get_uid(), population.get_uid(),
            [](size_t) { return STDPDeltaProjection::Synapse{{{1.0, 6, knp::synapse_traits::OutputType::EXCITATORY}, {1, 1}}, 0, 0}; },
            1};
        STDPDeltaProjection input{
            UID{false}, population.get_uid(),
            [](size_t) { return STDPDeltaProjection::Synapse{{{1.0, 1, knp::synapse_traits::OutputType::EXCITATORY}, {2, 2}}, 0, 0}; },
            1};
        loop.get_shared_parameters().stdp_populations_[population.get_uid()] =
            STDPDeltaProjection::SharedSynapseParameters::ProcessingType::STDPAndSpike;
        const UID input_uid = input.get_uid();
        backend.load_populations({population});
        backend.load_projections({input, loop});
        backend._init();
        CHECK(run_smallest(backend, input_uid, population.get_uid()) == kExpected);
        auto data = network_copy(backend);
        const auto &loop_after = std::get<STDPDeltaProjection>(data.second[1]);
        CHECK(std::get<knp::core::synapse_data>(loop_after[0]).weight_ != 1.0f);
    });

    run_case("ResourceSTDPBLIFATNetwork", [] { run_resource_network<knp::neuron_traits::SynapticResourceSTDPBLIFATNeuron>(); });
    run_case("ResourceSTDPAltAILIFNetwork", [] { run_resource_network<knp::neuron_traits::SynapticResourceSTDPAltAILIFNeuron>(); });

    run_case("LIFNetwork", [] {
        // one LIF neuron (leak 0.5, refractory period 2) fed 0.75 every step: it fires when the leaked
        // potential plus the impact exceeds 1 (steps 2, 6, 10, ...), then ignores two steps of impacts
        using LIFPopulation = knp::core::Population<knp::neuron_traits::LIFNeuron>;
        Backend backend;
        LIFPopulation population{[](size_t) {
                                     knp::neuron_traits::neuron_parameters<knp::neuron_traits::LIFNeuron> n;
                                     n.leak_coefficient_ = 0.5f;
                                     n.refract_period_ = 2;
                                     return n;
                                 },
                                 1};
        DeltaProjection input{UID{false}, population.get_uid(),
                              [](size_t) { return DeltaProjection::Synapse{{0.75, 1, knp::synapse_traits::OutputType::EXCITATORY}, 0, 0}; },
                              1};
        const UID input_uid = input.get_uid();
        backend.load_populations({population});
        backend.load_projections({input});
        backend._init();
        auto endpoint = backend.get_message_bus().create_endpoint();
        const UID in_channel_uid, out_channel_uid;
        backend.subscribe<SpikeMessage>(input_uid, {in_channel_uid});
        endpoint.subscribe<SpikeMessage>(out_channel_uid, {population.get_uid()});
        std::vector<knp::core::Step> results;
        for (knp::core::Step step = 0; step < 16; ++step) {
            endpoint.send_message(SpikeMessage{{in_channel_uid, 0}, {0}});
            backend._step();
            endpoint.receive_all_messages();
            if (!endpoint.unload_messages<SpikeMessage>(out_channel_uid).empty()) results.push_back(step);
        }
        CHECK((results == std::vector<knp::core::Step>{2, 6, 10, 14}));
        const auto data = network_copy(backend);
        const auto &after = std::get<LIFPopulation>(data.first[0]);
        CHECK(after[0].refract_counter_ == 1);  // fired at step 14, one refractory step done
        CHECK(after[0].potential_ == 0.0f);
    });

    run_case("AltAILIFNetwork", [] {
        // SpikeGenerationTest of knp/tests/framework/altai_lif_test.cpp:162-175: leak 50, threshold 110,
        // no input: the potential climbs 0, 50, 100 and fires on steps 2, 5, 8
        using AltAIPopulation = knp::core::Population<knp::neuron_traits::AltAILIF>;
        Backend backend;
        AltAIPopulation population{[](size_t) {
                                       knp::neuron_traits::neuron_parameters<knp::neuron_traits::AltAILIF> n;
                                       n.potential_leak_ = 50;
                                       n.activation_threshold_ = 110;
                                       return n;
                                   },
                                   1};
        backend.load_populations({population});
        backend.load_projections({});
        backend._init();
        auto endpoint = backend.get_message_bus().create_endpoint();
        const UID out_channel_uid;
        endpoint.subscribe<SpikeMessage>(out_channel_uid, {population.get_uid()});
        std::vector<knp::core::Step> results;
        std::vector<float> potentials;
        for (knp::core::Step step = 0; step < 10; ++step) {
            potentials.push_back(std::get<AltAIPopulation>(network_copy(backend).first[0])[0].potential_);
            backend._step();
            endpoint.receive_all_messages();
            if (!endpoint.unload_messages<SpikeMessage>(out_channel_uid).empty()) results.push_back(step);
        }
        CHECK((results == std::vector<knp::core::Step>{2, 5, 8}));
        CHECK((potentials == std::vector<float>{0, 50, 100, 0, 50, 100, 0, 50, 100, 0}));
    });

    run_case("NeuronsAndSynapsesGetting", [] {
        Backend backend;
        CHECK(backend.get_supported_neurons()[0] == "BLIFATNeuron");
        CHECK(backend.get_supported_neurons().back() == "LIFNeuron");
        CHECK(backend.get_supported_synapses()[0] == "DeltaSynapse");
        CHECK(backend.plasticity_supported());
        CHECK((backend.get_supported_population_indexes() == std::vector<size_t>{0, 1, 2, 3, 4}));
        CHECK((backend.get_supported_projection_indexes() == std::vector<size_t>{0, 1, 2}));
    });

    run_case("NetworkDataThroughConstBackend", [] {
        // get_network_data through a const Backend (as get_network_copy and the weights logger call it): the
        // ranges yield copies with the device state, and a second call sees the state of later steps
        Backend backend;
        BLIFATPopulation population{[](size_t) { return knp::neuron_traits::neuron_parameters<knp::neuron_traits::BLIFATNeuron>{}; }, 3};
        DeltaProjection input{UID{false}, population.get_uid(),
                              [](size_t i) { return DeltaProjection::Synapse{{0.25f * float(i + 1), 1, knp::synapse_traits::OutputType::EXCITATORY}, i, i}; },
                              3};
        const UID input_uid = input.get_uid();
        backend.load_populations({population});
        backend.load_projections({input});
        backend._init();
        const knp::core::Backend &cb = backend;
        auto endpoint = backend.get_message_bus().create_endpoint();
        const UID in_channel_uid;
        backend.subscribe<SpikeMessage>(input_uid, {in_channel_uid});
        endpoint.send_message(SpikeMessage{{in_channel_uid, 0}, {0, 1, 2}});
        backend._step();
        backend._step();
        auto net = network_copy(cb);
        CHECK(net.first.size() == 1 && net.second.size() == 1);
        const auto &p = std::get<BLIFATPopulation>(net.first[0]);
        CHECK(p.size() == 3 && p.get_uid() == population.get_uid());
        // impacts 0.25, 0.5, 0.75 arrived at step 1 below the threshold 1: the potentials hold them
        CHECK(p[0].potential_ == 0.25 && p[1].potential_ == 0.5 && p[2].potential_ == 0.75);
        const auto &q = std::get<DeltaProjection>(net.second[0]);
        CHECK(q.size() == 3 && std::get<knp::core::synapse_data>(q[2]).weight_ == 0.75f);
        backend._step();
        auto later = network_copy(cb);
        CHECK(std::get<BLIFATPopulation>(later.first[0])[1].potential_ == 0.0);  // memoryless: no input at step 2
        // two ranges from one call stay independent of later steps
        auto ranges = cb.get_network_data();
        CHECK(*ranges.population_range.first != *ranges.population_range.second);
    });

    run_case("SelectDevice", [] {
        Backend backend;
        auto devices = backend.get_devices();
        CHECK(!devices.empty());
        const UID uid = devices[0]->get_uid();
        backend.select_device(std::move(devices[0]));
        CHECK(backend.get_current_devices().size() == 1 && backend.get_current_devices()[0]->get_uid() == uid);
        backend.select_devices({uid});
        CHECK(!backend.get_current_devices().empty());
    });

    run_case("BackendLoader", [&] {
        knp::framework::BackendLoader loader;
        CHECK(knp::framework::BackendLoader::is_backend(lib));
        auto backend = loader.load(lib);
        CHECK(backend != nullptr);
        CHECK(backend->get_supported_neurons()[0] == "BLIFATNeuron");
        CHECK(!backend->get_devices().empty());
        bool threw = false;
        try {
            loader.load(lib.parent_path() / "no-such-backend");
        } catch (const std::runtime_error &) {
            threw = true;
        }
        CHECK(threw);
        threw = false;
        try {
            backend->select_devices({UID()});
        } catch (const std::logic_error &) {
            threw = true;
        }
        CHECK(threw);
    });

    // MultiThreadCpuSuite SmallestNetwork and SmallestResourceNetwork (multi_thread_cpu_test.cpp:54-138) through
    // the multi-threaded plugin
    run_case("MultiThreadedSmallestNetwork", [&] {
        knp::framework::BackendLoader loader;
        auto backend = loader.load(mt_lib);
        BLIFATPopulation population{[](size_t) { return knp::neuron_traits::neuron_parameters<knp::neuron_traits::BLIFATNeuron>{}; }, 1};
        DeltaProjection loop{population.get_uid(), population.get_uid(),
                             [](size_t) { return DeltaProjection::Synapse{{1.1, 6, knp::synapse_traits::OutputType::EXCITATORY}, 0, 0}; },
                             1};
        DeltaProjection input{UID{false}, population.get_uid(),
                              [](size_t) { return DeltaProjection::Synapse{{1.0, 1, knp::synapse_traits::OutputType::EXCITATORY}, 0, 0}; },
                              1};
        const UID input_uid = input.get_uid();
        backend->load_all_populations({population});
        backend->load_all_projections({input, loop});
        CHECK(run_smallest(*backend, input_uid, population.get_uid()) == kExpected);
    });

    run_case("MultiThreadedSmallestResourceNetwork", [&] {
        using STDPDeltaProjection = knp::core::Projection<knp::synapse_traits::SynapticResourceSTDPDeltaSynapse>;
        using ResourcePopulation = knp::core::Population<knp::neuron_traits::SynapticResourceSTDPBLIFATNeuron>;
        knp::framework::BackendLoader loader;
        auto backend = loader.load(mt_lib);
        ResourcePopulation population{[](size_t) { return knp::neuron_traits::neuron_parameters<knp::neuron_traits::SynapticResourceSTDPBLIFATNeuron>{}; }, 1};
        auto synapse = [](uint32_t delay) {
            return [delay](size_t) {
                knp::synapse_traits::synapse_parameters<knp::synapse_traits::SynapticResourceSTDPDeltaSynapse> syn;
                syn.rule_.w_max_ = 2.0;
                syn.rule_.w_min_ = 1.0;
                syn.rule_.synaptic_resource_ = 1.0;
                syn.rule_.dopamine_plasticity_period_ = 5;
                syn.weight_ = 1.5;
                syn.delay_ = delay;
                syn.output_type_ = knp::synapse_traits::OutputType::EXCITATORY;
                return STDPDeltaProjection::Synapse{syn, 0, 0};
            };
        };
        STDPDeltaProjection loop{population.get_uid(), population.get_uid(), synapse(6), 1};
        STDPDeltaProjection input{UID{false}, population.get_uid(), synapse(1), 1};
        const UID input_uid = input.get_uid();
        backend->load_all_populations({population});
        backend->load_all_projections({input, loop});
        CHECK(run_smallest(*backend, input_uid, population.get_uid()) == kExpected);
    });

    run_case("MultiThreadedRepeatedIndex", [&] {
        knp::framework::BackendLoader loader;
        auto mt = loader.load(mt_lib);
        CHECK((repeated_index_spikes(*mt) == std::vector<uint32_t>{0}));
        auto st = loader.load(lib);
        CHECK(repeated_index_spikes(*st).empty());
    });

    run_case("StartWithStopPredicate", [&] {
        knp::framework::BackendLoader loader;
        auto backend = loader.load(lib);
        BLIFATPopulation population{[](size_t) { return knp::neuron_traits::neuron_parameters<knp::neuron_traits::BLIFATNeuron>{}; }, 1};
        DeltaProjection loop{population.get_uid(), population.get_uid(),
                             [](size_t) { return DeltaProjection::Synapse{{1.1, 6, knp::synapse_traits::OutputType::EXCITATORY}, 0, 0}; },
                             1};
        backend->load_all_populations({population});
        backend->load_all_projections({loop});
        backend->start([](knp::core::Step s) { return s < 7; });
        CHECK(backend->get_step() == 7);
        CHECK(!backend->running());
    });

    return g_failures ? 1 : 0;
}
