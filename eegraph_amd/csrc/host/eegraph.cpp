// Graph loading, the Tigr virtual graph handle, the eegraph_<alg> drivers, utilities, timer
// and the energy sampler, on top of the C-ABI.  Follows src/graph.cpp, src/virtual_graph.cpp,
// src/eegraph.cu (driver steps: build, init, run, copy back, optional CPU check, save),
// src/utilities.cpp and src/timer.cpp of the reference.
#include <algorithm>
#include <cstdio>
#include <deque>
#include <fstream>
#include <iostream>
#include <numeric>

#include "eegraph_host.hpp"

// ------------------------------------------------------------------------------------ Graph
Graph::Graph(std::string graphFilePath, bool isWeighted)
    : graphFilePath(std::move(graphFilePath)), isWeighted(isWeighted) {}

std::string Graph::getFileExtension(std::string fileName) {
    const auto dot = fileName.find_last_of('.');
    return dot == std::string::npos ? std::string() : fileName.substr(dot + 1);
}

bool Graph::ReadGraph() {
    std::cout << "Reading the input graph from the following file:\n>> " << graphFilePath << std::endl;
    graphFormat = getFileExtension(graphFilePath);
    eeg_edgelist el;
    const eeg_status s = eeg_load_edgelist(graphFilePath.c_str(), isWeighted ? 1 : 0, &el);
    if (s != EEG_OK) {
        std::cout << "Graph file could not be read: " << eeg_strerror(s) << " (" << eeg_last_error()
                  << ")" << std::endl;
        return false;
    }
    edges.resize(el.m);
    for (uint64_t i = 0; i < el.m; ++i) edges[i] = Edge{el.src[i], el.dst[i]};
    weights.assign(isWeighted ? el.m : 0, 1u);
    if (isWeighted && el.w)
        for (uint64_t i = 0; i < el.m; ++i) weights[i] = el.w[i];
    num_nodes = el.n;
    num_edges = el.m;
    hasZeroID = el.has_zero_id != 0;
    graphLoaded = true;
    eeg_edgelist_free(&el);
    std::cout << "Done reading.\n";
    std::cout << "Number of nodes = " << num_nodes << std::endl;
    std::cout << "Number of edges = " << num_edges << std::endl;
    return true;
}

// ----------------------------------------------------------------------------- VirtualGraph
VirtualGraph::VirtualGraph(Graph& g, uint K, int deviceID) : graph(&g), K(K), deviceID(deviceID) {
    // in/out degrees (src/virtual_graph.cpp:16-32), without touching the caller's edges
    inDegree = new uint[g.num_nodes ? g.num_nodes : 1]();
    outDegree = new uint[g.num_nodes ? g.num_nodes : 1]();
    for (const Edge& e : g.edges) {
        outDegree[e.source]++;
        inDegree[e.end]++;
    }
}

VirtualGraph::~VirtualGraph() {
    if (device) eeg_graph_destroy(device);
    delete[] inDegree;
    delete[] outDegree;
    delete[] nodePointer;
    delete[] edgeList;
    delete[] partNodePointer;
}

eeg_status VirtualGraph::MakeGraph() {
    const Graph& g = *graph;
    std::vector<uint32_t> s(g.num_edges), d(g.num_edges);
    for (uint64_t i = 0; i < g.num_edges; ++i) {
        s[i] = g.edges[i].source;
        d[i] = g.edges[i].end;
    }
    const bool wtd = g.isWeighted && g.weights.size() == g.num_edges;
    weights = wtd ? g.weights : std::vector<uint>();
    if (device) {
        eeg_graph_destroy(device);
        device = nullptr;
    }
    eeg_status st = eeg_graph_create(s.data(), d.data(), wtd ? g.weights.data() : nullptr,
                                     g.num_edges, g.num_nodes, deviceID, &device);
    if (st != EEG_OK) return st;
    st = eeg_graph_prepare(device, K, EEG_PREP_TIGR);
    if (st != EEG_OK) return st;
    return eeg_tigr_export(device, K, nullptr, nullptr, nullptr, &numParts);
}

eeg_status VirtualGraph::ExportLayout() {
    if (!device) return EEG_ERR_INVALID_ARG;
    const uint64_t n = graph->num_nodes, m = graph->num_edges;
    delete[] nodePointer;
    delete[] edgeList;
    delete[] partNodePointer;
    nodePointer = new uint64_t[n ? n : 1];
    edgeList = new uint[2 * m + n + 1];
    partNodePointer = new PartPointer[numParts ? numParts : 1];
    return eeg_tigr_export(device, K, nodePointer, edgeList, (uint32_t*)partNodePointer, &numParts);
}

// ---------------------------------------------------------------------------------- drivers
namespace {
const char* algo_name(uint32_t a) {
    static const char* names[] = {"bfs", "sssp", "sswp", "cc", "pr", "ppr"};
    return a <= EEG_PPR ? names[a] : "?";
}
}  // namespace

Result eegraph_run(uint32_t algo, ArgumentParser& arguments, Graph& graph) {
    Result result{};
    result.time = 0;
    result.energy = 0;
    EnergySampler nvml(arguments.deviceID, arguments.energyFile, arguments.energyStats);
    if (arguments.energy) {
        std::cout << "Starting energy measurements..." << std::endl;
        if (!nvml.start()) std::cout << "amd-smi energy counter unavailable; energy reported as 0" << std::endl;
    }
    if (arguments.unifiedMem)
        std::cout << "Note: --um is accepted for compatibility; the graph lives in HBM" << std::endl;

    VirtualGraph vGraph(graph, arguments.K, arguments.deviceID);
    eeg_status st = vGraph.MakeGraph();
    if (st == EEG_OK) {
        uint32_t what = EEG_PREP_TIGR;
        if (algo == EEG_CC) what |= EEG_PREP_SYM;
        if (algo == EEG_PR) what |= EEG_PREP_TRANSPOSE;
        st = eeg_graph_prepare(vGraph.device, arguments.K, what);
    }
    if (st != EEG_OK) {
        std::cout << "Graph construction failed: " << eeg_strerror(st) << " (" << eeg_last_error() << ")"
                  << std::endl;
        result.status = st;
        nvml.stop();
        return result;
    }
    const uint num_nodes = graph.num_nodes;
    const bool sourced = (algo == EEG_BFS || algo == EEG_SSSP || algo == EEG_SSWP);
    if (sourced && (arguments.sourceNode < 0 || (uint)arguments.sourceNode >= num_nodes)) {
        std::cout << "Source " << arguments.sourceNode << " is out of range (n = " << num_nodes << ")" << std::endl;
        result.status = EEG_ERR_INVALID_ARG;
        nvml.stop();
        return result;
    }

    eeg_opts o{};
    o.variant = arguments.variant;
    o.K = arguments.K;
    o.frontier = arguments.frontierKind;
    o.wgs_per_cu = arguments.wgsPerCU;
    o.collect_activity = arguments.nodeActivity ? 1 : 0;
    if (algo == EEG_PR && arguments.hasAcc) o.epsilon = arguments.acc;
    if (arguments.autotune) {
        uint32_t sample = sourced ? (uint32_t)arguments.sourceNode : 0;
        double best_ms = 0;
        eeg_opts tuned = o;
        if (eeg_autotune(vGraph.device, algo, &sample, 1, &tuned, &best_ms) == EEG_OK) {
            tuned.collect_activity = o.collect_activity;
            o = tuned;
            std::cout << "Autotuned: K=" << o.K << " variant=" << o.variant << " wgs=" << o.wgs_per_cu
                      << " frontier=" << o.frontier << " (" << best_ms << " ms)" << std::endl;
        }
    }

    std::vector<uint32_t> labels(num_nodes);
    std::vector<uint32_t> activity(arguments.nodeActivity ? 1u << 20 : 0);
    eeg_stats stats{};
    stats.activity = activity.empty() ? nullptr : activity.data();
    stats.activity_cap = (uint32_t)activity.size();

    Timer totalTimer;
    totalTimer.Start();
    if (arguments.energy) nvml.log_point();
    st = eeg_run(vGraph.device, algo, sourced ? (uint32_t)arguments.sourceNode : 0, &o, labels.data(), &stats);
    if (arguments.energy) nvml.log_point();
    const float total = totalTimer.Finish();
    nvml.stop();

    result.status = st;
    if (st != EEG_OK && st != EEG_ERR_NOT_CONVERGED) {
        std::cout << algo_name(algo) << " failed: " << eeg_strerror(st) << " (" << eeg_last_error() << ")" << std::endl;
        return result;
    }
    std::cout << "Number of iterations = " << stats.iterations << std::endl;
    std::cout << "Processing finished in " << stats.ms_traversal << " (ms).\n";
    std::cout << "Total GPU activity finished in " << total << " (ms).\n";
    result.time = total;
    result.energy = nvml.energy_mJ();
    result.iterations = stats.iterations;
    result.traversal_ms = stats.ms_traversal;
    result.sumLabelsVec.assign(activity.begin(), activity.begin() + stats.activity_len);
    if (arguments.energy) std::cout << "Energy = " << result.energy << " (mJ)" << std::endl;

    if (arguments.debug) {
        if (algo == EEG_PR) {
            utilities::PrintResults((float*)labels.data(), std::min<uint>(30, num_nodes));
        } else {
            std::vector<uint> cpu;
            utilities::CpuLabels(algo, graph, (uint)arguments.sourceNode, cpu);
            utilities::PrintResults(labels.data(), std::min<uint>(30, num_nodes));
            const uint64_t bad = utilities::CompareArrays(cpu.data(), labels.data(), num_nodes);
            std::cout << (bad ? "CPU check: MISMATCH" : "CPU check: identical") << " (" << bad
                      << " differing labels)" << std::endl;
        }
    }
    if (arguments.hasOutput) {
        if (algo == EEG_PR) utilities::SaveResults(arguments.output, (float*)labels.data(), num_nodes);
        else utilities::SaveResults(arguments.output, labels.data(), num_nodes);
    }
    return result;
}

Result eegraph_bfs(ArgumentParser& a, Graph& g) { return eegraph_run(EEG_BFS, a, g); }
Result eegraph_sssp(ArgumentParser& a, Graph& g) { return eegraph_run(EEG_SSSP, a, g); }
Result eegraph_sswp(ArgumentParser& a, Graph& g) { return eegraph_run(EEG_SSWP, a, g); }
Result eegraph_cc(ArgumentParser& a, Graph& g) { return eegraph_run(EEG_CC, a, g); }
Result eegraph_pr(ArgumentParser& a, Graph& g) { return eegraph_run(EEG_PR, a, g); }
Result eegraph_bfs_um(ArgumentParser& a, Graph& g) { return eegraph_run(EEG_BFS, a, g); }
Result eegraph_sssp_um(ArgumentParser& a, Graph& g) { return eegraph_run(EEG_SSSP, a, g); }
Result eegraph_sswp_um(ArgumentParser& a, Graph& g) { return eegraph_run(EEG_SSWP, a, g); }
Result eegraph_cc_um(ArgumentParser& a, Graph& g) { return eegraph_run(EEG_CC, a, g); }
Result eegraph_pr_um(ArgumentParser& a, Graph& g) { return eegraph_run(EEG_PR, a, g); }

// -------------------------------------------------------------------------------- utilities
namespace utilities {

void PrintResults(uint* results, uint n) {
    std::cout << "Results of first " << n << " nodes:\n[";
    for (uint i = 0; i < n; ++i) std::cout << (i ? " " : "") << i << ":" << results[i];
    std::cout << "]\n";
}

void PrintResults(float* results, uint n) {
    std::cout << "Results of first " << n << " nodes:\n[";
    for (uint i = 0; i < n; ++i) std::cout << (i ? " " : "") << i << ":" << results[i];
    std::cout << "]\n";
}

void SaveResults(std::string filepath, uint* results, uint n) {
    std::cout << "Saving the results into the following file:\n>> " << filepath << std::endl;
    std::ofstream out(filepath);
    for (uint i = 0; i < n; ++i) out << i << " " << results[i] << "\n";
    std::cout << "Done saving.\n";
}

void SaveResults(std::string filepath, float* results, uint n) {
    std::cout << "Saving the results into " << filepath << " ...... " << std::flush;
    std::ofstream out(filepath);
    out.precision(9);
    for (uint i = 0; i < n; ++i) out << i << " " << results[i] << "\n";
    std::cout << " Completed.\n";
}

uint64_t CompareArrays(unsigned int* arr1, unsigned int* arr2, uint n) {
    uint64_t bad = 0;
    for (uint i = 0; i < n; ++i) {
        if (arr1[i] != arr2[i]) {
            if (!bad) {
                std::cout << "Arrays begin to differ at element " << i << std::endl;
                std::cout << "Arr1[" << i << "]: " << arr1[i] << std::endl;
                std::cout << "Arr2[" << i << "]: " << arr2[i] << std::endl;
            }
            ++bad;
        }
    }
    return bad;
}

double maxActivePct(std::vector<unsigned int> activeNodes, unsigned int total) {
    unsigned int mx = 0;
    for (unsigned int a : activeNodes) mx = std::max(mx, a);
    return total ? (double)mx / (double)total : 0.0;
}

double pctIterOverThreshold(std::vector<unsigned int> activeNodes, unsigned int total, double threshold) {
    if (activeNodes.empty() || !total) return 0.0;
    unsigned int count = 0;
    for (unsigned int a : activeNodes)
        if ((double)a / (double)total > threshold) ++count;
    return (double)count / (double)activeNodes.size();
}

void CpuLabels(uint32_t algo, const Graph& g, uint source, std::vector<uint>& out) {
    const uint n = g.num_nodes;
    out.assign(n, 0);
    if (algo == EEG_CC) {  // union-find, label = smallest id of the component
        std::vector<uint> p(n);
        std::iota(p.begin(), p.end(), 0u);
        auto find = [&](uint x) {
            while (p[x] != x) x = p[x] = p[p[x]];
            return x;
        };
        for (const Edge& e : g.edges) {
            uint a = find(e.source), b = find(e.end);
            if (a != b) (a < b ? p[b] : p[a]) = (a < b ? a : b);
        }
        for (uint v = 0; v < n; ++v) out[v] = find(v);
        return;
    }
    std::vector<uint64_t> row(n + 1, 0);
    for (const Edge& e : g.edges) row[e.source + 1]++;
    for (uint v = 0; v < n; ++v) row[v + 1] += row[v];
    std::vector<uint> col(g.num_edges), wt(g.num_edges);
    std::vector<uint64_t> cur(row.begin(), row.end() - 1);
    for (uint64_t i = 0; i < g.num_edges; ++i) {
        const uint64_t k = cur[g.edges[i].source]++;
        col[k] = g.edges[i].end;
        wt[k] = g.weights.empty() ? 1u : g.weights[i];
    }
    const bool widest = (algo == EEG_SSWP);
    for (uint v = 0; v < n; ++v) out[v] = widest ? 0u : DIST_INFINITY;
    out[source] = widest ? DIST_INFINITY : 0u;
    std::deque<uint> q{source};
    std::vector<char> inq(n, 0);
    inq[source] = 1;
    while (!q.empty()) {  // label-correcting FIFO relaxation
        const uint u = q.front();
        q.pop_front();
        inq[u] = 0;
        for (uint64_t e = row[u]; e < row[u + 1]; ++e) {
            const uint v = col[e];
            uint c;
            if (widest) c = std::min(out[u], wt[e]);
            else if (algo == EEG_BFS) c = out[u] + 1;
            else c = (uint)std::min<uint64_t>((uint64_t)out[u] + wt[e], DIST_INFINITY);
            const bool better = widest ? c > out[v] : c < out[v];
            if (better) {
                out[v] = c;
                if (!inq[v]) {
                    inq[v] = 1;
                    q.push_back(v);
                }
            }
        }
    }
}

}  // namespace utilities

// ------------------------------------------------------------------------------------ Timer
void Timer::Start() { start_ = std::chrono::steady_clock::now(); }
float Timer::Finish() {
    return std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - start_).count();
}

// --------------------------------------------------------------------------- EnergySampler
EnergySampler::EnergySampler(int device, std::string readingsFile, std::string statsFile)
    : device_(device), readings_(std::move(readingsFile)), stats_(std::move(statsFile)) {}

EnergySampler::~EnergySampler() { stop(); }

bool EnergySampler::start() {
    double uj = 0;
    if (eeg_energy_read(device_, &uj, nullptr) != EEG_OK) return false;
    t0_ = std::chrono::steady_clock::now();
    running_ = true;
    th_ = std::thread(&EnergySampler::loop, this);
    return true;
}

void EnergySampler::loop() {
    while (running_.load()) {
        double uj = 0;
        if (eeg_energy_read(device_, &uj, nullptr) == EEG_OK) {
            const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
            samples_.push_back({t, uj});
        }
        std::this_thread::sleep_for(std::chrono::milliseconds(1));
    }
}

void EnergySampler::log_point() {
    double uj = 0;
    if (eeg_energy_read(device_, &uj, nullptr) != EEG_OK) return;
    const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
    points_.push_back({t, uj});
}

void EnergySampler::stop() {
    if (!running_.exchange(false)) return;
    if (th_.joinable()) th_.join();
    if (!readings_.empty()) {
        std::ofstream f(readings_);
        f << "time_ms,energy_uj,power_w\n";
        for (size_t i = 0; i < samples_.size(); ++i) {
            double pw = 0;
            if (i) {
                const double dt = samples_[i].t_ms - samples_[i - 1].t_ms;
                pw = dt > 0 ? (samples_[i].uj - samples_[i - 1].uj) / (dt * 1e3) : 0;
            }
            f << samples_[i].t_ms << "," << samples_[i].uj << "," << pw << "\n";
        }
    }
    if (!stats_.empty()) {
        std::ofstream f(stats_);
        f << "point,time_ms,energy_uj\n";
        for (size_t i = 0; i < points_.size(); ++i)
            f << i << "," << points_[i].t_ms << "," << points_[i].uj << "\n";
        f << "energy_mJ," << energy_mJ() << "\n";
    }
}

double EnergySampler::energy_mJ() const {
    if (points_.size() < 2) return 0.0;
    return (points_.back().uj - points_.front().uj) * 1e-3;
}
