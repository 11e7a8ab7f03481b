// C++ driver surface of the reference (argument parsing, Graph / VirtualGraph, the
// eegraph_<alg> entry points, utilities, timer, energy), rebuilt on top of the C-ABI in
// include/eeg_capi.h.  Host code only: the GPU is reached through libeegraph.so.
//
// Mirrors include/globals.hpp, include/graph.hpp, include/virtual_graph.hpp,
// include/argument_parsing.hpp, include/eegraph.cuh, include/utilities.hpp, include/timer.hpp
// of the reference; differences are noted where they occur.
#pragma once

#include <atomic>
#include <chrono>
#include <climits>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/eeg_capi.h"

typedef unsigned int uint;

// include/globals.hpp:30-57.  Part_Size is the default K; the virtual-node bound is a runtime
// parameter here (--k).
const unsigned int Part_Size = 8;
const unsigned int DIST_INFINITY = UINT_MAX - 1;

enum Variant : unsigned char {  // same order as the reference (globals.hpp:36-41)
    ASYNC_PUSH_TD,
    ASYNC_PUSH_DD,
    SYNC_PUSH_TD,
    SYNC_PUSH_DD,
};

struct Edge {
    unsigned int source;
    unsigned int end;
};

struct PartPointer {
    unsigned int node;
    unsigned int part;
};

struct Result {
    float time;                              // total ms, as the reference (eegraph.cu:2236)
    double energy;                           // mJ over the run (0 without --energy)
    std::vector<unsigned int> sumLabelsVec;  // per-iteration frontier sizes (--activity)
    // additions
    unsigned int iterations = 0;
    double traversal_ms = 0;                 // device events, first kernel to convergence
    int status = 0;                          // eeg_status of the run
};

// src/argument_parsing.cpp: the same flag pairs and defaults; errors print the help text and
// exit(0) like the reference.  Added flags: --k, --wgs, --frontier, --autotune.
class ArgumentParser {
public:
    int argc;
    char** argv;

    bool canHaveSource;
    bool canHaveItrs;

    bool hasInput = false;
    bool hasSourceNode = false;
    bool hasOutput = false;
    bool hasDeviceID = false;
    bool hasNumberOfItrs = false;
    bool hasEnergyFile = false;
    bool hasEnergyStats = false;
    bool hasAcc = false;

    bool debug = false;
    bool energy = false;
    bool unifiedMem = false;
    bool nodeActivity = false;

    std::string input;
    std::string output;
    std::string energyFile;
    std::string energyStats;

    int sourceNode = 0;
    int deviceID = 0;
    int numberOfItrs = 1;

    float acc = 0.01f;

    Variant variant = ASYNC_PUSH_TD;

    // additions
    unsigned int K = Part_Size;
    unsigned int wgsPerCU = 0;
    unsigned int frontierKind = 0;  // eeg_frontier_kind of sync_push_dd (0 = sorted)
    bool autotune = false;
    std::string algorithm;  // for the combined `eegraph` executable

    ArgumentParser(int argc, char** argv, bool canHaveSource, bool canHaveItrs);
    bool Parse();
    std::string GenerateHelpString();
};

// src/graph.cpp: text edge lists (.el .wel .txt .edges .edgelist) with the reference's parsing
// rules, plus Subway's binary .bcsr/.bwcsr.  Ids are normalised to 0-based here, once; the
// reference shifts the caller's edges inside every VirtualGraph constructor (defect D8).
class Graph {
public:
    std::string graphFilePath;
    std::string graphFormat;
    bool isWeighted;
    bool hasZeroID = false;
    bool graphLoaded = false;
    uint num_nodes = 0;
    uint64_t num_edges = 0;
    std::vector<Edge> edges;
    std::vector<uint> weights;

    Graph(std::string graphFilePath, bool isWeighted);
    bool ReadGraph();  // false (and a message) when the file cannot be read
    std::string getFileExtension(std::string fileName);
};

// src/um_graph.cu / include/um_graph.cuh: the unified-memory twin of Graph, with the same
// constructor and ReadGraph().  Graphs live in HBM here, so it is the same class and the
// eegraph_<alg>_um entry points below take it (src/main.cu:45,309-325 compile unchanged).
using UMGraph = Graph;

// src/virtual_graph.cpp: the Tigr virtual graph.  MakeGraph() uploads the edges and builds the
// CSR and the virtual-node tables in HBM; the host arrays of the reference layout
// (nodePointer, edgeList, partNodePointer) are filled only by ExportLayout(), for inspection.
class VirtualGraph {
public:
    Graph* graph;
    eeg_graph* device = nullptr;
    uint K;
    int deviceID;
    // reference layout (ExportLayout)
    uint64_t* nodePointer = nullptr;  // reference: uint (overflows past 2^32 slots)
    uint* edgeList = nullptr;
    uint* inDegree = nullptr;
    uint* outDegree = nullptr;
    uint64_t numParts = 0;
    PartPointer* partNodePointer = nullptr;
    std::vector<uint> weights;

    VirtualGraph(Graph& graph, uint K = Part_Size, int deviceID = 0);
    ~VirtualGraph();
    VirtualGraph(const VirtualGraph&) = delete;
    VirtualGraph& operator=(const VirtualGraph&) = delete;
    eeg_status MakeGraph();
    eeg_status ExportLayout();
};

// include/eegraph.cuh:21-39.  The _um forms take a UMGraph (= Graph) and run the same
// HBM-resident path (the unified-memory build is replaced by an explicit device CSR).
Result eegraph_bfs(ArgumentParser& arguments, Graph& graph);
Result eegraph_sssp(ArgumentParser& arguments, Graph& graph);
Result eegraph_sswp(ArgumentParser& arguments, Graph& graph);
Result eegraph_cc(ArgumentParser& arguments, Graph& graph);
Result eegraph_pr(ArgumentParser& arguments, Graph& graph);
Result eegraph_bfs_um(ArgumentParser& arguments, UMGraph& graph);
Result eegraph_sssp_um(ArgumentParser& arguments, UMGraph& graph);
Result eegraph_sswp_um(ArgumentParser& arguments, UMGraph& graph);
Result eegraph_cc_um(ArgumentParser& arguments, UMGraph& graph);
Result eegraph_pr_um(ArgumentParser& arguments, UMGraph& graph);

// One driver for all algorithms (what each *_main.cu duplicates in the reference).
Result eegraph_run(uint32_t algo, ArgumentParser& arguments, Graph& graph);

namespace utilities {  // src/utilities.cpp
void PrintResults(uint* results, uint n);
void PrintResults(float* results, uint n);
void SaveResults(std::string filepath, uint* results, uint n);
void SaveResults(std::string filepath, float* results, uint n);
// Prints the first mismatch like the reference and returns the number of mismatches.
uint64_t CompareArrays(unsigned int* arr1, unsigned int* arr2, uint n);
double maxActivePct(std::vector<unsigned int> activeNodes, unsigned int total);
double pctIterOverThreshold(std::vector<unsigned int> activeNodes, unsigned int total, double threshold);
// Serial label-correcting CPU run of BFS/SSSP/SSWP/CC used by --debug true (the reference runs
// its seq_cpu sweep there, src/eegraph.cu:239-257).
void CpuLabels(uint32_t algo, const Graph& graph, uint source, std::vector<uint>& out);
}  // namespace utilities

// src/timer.cpp (steady clock instead of gettimeofday)
class Timer {
    std::chrono::steady_clock::time_point start_;
public:
    void Start();
    float Finish();  // ms since Start
};

// Energy side path (replaces include/nvmlClass.cuh): a sampler thread reads the amd-smi energy
// accumulator every millisecond; log_point() marks phases; energy between the first and last
// marks is the difference of two counter reads.
class EnergySampler {
public:
    EnergySampler(int device, std::string readingsFile, std::string statsFile);
    ~EnergySampler();
    bool start();       // false if the counter is unavailable
    void log_point();
    void stop();
    double energy_mJ() const;  // between the first and the last log_point
private:
    int device_;
    std::string readings_, stats_;
    std::atomic<bool> running_{false};
    std::thread th_;
    struct Sample {
        double t_ms;
        double uj;
    };
    std::vector<Sample> samples_;
    std::vector<Sample> points_;
    std::chrono::steady_clock::time_point t0_;
    void loop();
};
