// Command-line parsing with the reference's flags, defaults and error behaviour
// (src/argument_parsing.cpp:3-175): flag/value pairs, help text and exit(0) on any error.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "eegraph_host.hpp"

namespace {
bool is_true(const char* v) {
    return strcmp(v, "true") == 0 || strcmp(v, "True") == 0 || strcmp(v, "TRUE") == 0;
}
}  // namespace

ArgumentParser::ArgumentParser(int argc, char** argv, bool canHaveSource, bool canHaveItrs)
    : argc(argc), argv(argv), canHaveSource(canHaveSource), canHaveItrs(canHaveItrs) {
    Parse();
}

bool ArgumentParser::Parse() {
    if (argc == 1) {
        std::cout << GenerateHelpString();
        exit(0);
    }
    if (argc == 2 && (strcmp(argv[1], "--help") == 0 || strcmp(argv[1], "-help") == 0 ||
                      strcmp(argv[1], "--h") == 0 || strcmp(argv[1], "-h") == 0)) {
        std::cout << GenerateHelpString();
        exit(0);
    }
    if (argc % 2 == 0) {
        std::cout << "\nThere was an error parsing command line arguments\n";
        std::cout << GenerateHelpString();
        exit(0);
    }
    for (int i = 1; i < argc - 1; i += 2) {
        const char* f = argv[i];
        const char* v = argv[i + 1];
        if (strcmp(f, "--input") == 0) {
            input = v;
            hasInput = true;
        } else if (strcmp(f, "--output") == 0) {
            output = v;
            hasOutput = true;
        } else if (strcmp(f, "--efile") == 0) {
            energyFile = v;
            hasEnergyFile = true;
        } else if (strcmp(f, "--estats") == 0) {
            energyStats = v;
            hasEnergyStats = true;
        } else if (strcmp(f, "--source") == 0 && canHaveSource) {
            sourceNode = atoi(v);
            hasSourceNode = true;
        } else if (strcmp(f, "--device") == 0) {
            deviceID = atoi(v);
            hasDeviceID = true;
        } else if (strcmp(f, "--iteration") == 0 && canHaveItrs) {
            numberOfItrs = atoi(v);
            hasNumberOfItrs = true;
        } else if (strcmp(f, "--debug") == 0) {
            if (is_true(v)) debug = true;
        } else if (strcmp(f, "--activity") == 0) {
            if (is_true(v)) nodeActivity = true;
        } else if (strcmp(f, "--variant") == 0) {
            if (strcmp(v, "async_push_td") == 0) variant = ASYNC_PUSH_TD;
            else if (strcmp(v, "async_push_dd") == 0) variant = ASYNC_PUSH_DD;
            else if (strcmp(v, "sync_push_td") == 0) variant = SYNC_PUSH_TD;
            else if (strcmp(v, "sync_push_dd") == 0) variant = SYNC_PUSH_DD;
            else {
                std::cout << "Variant not recognized\n";
                exit(0);
            }
        } else if (strcmp(f, "--energy") == 0) {
            if (is_true(v)) energy = true;
        } else if (strcmp(f, "--um") == 0) {
            if (is_true(v)) unifiedMem = true;
        } else if (strcmp(f, "--accuracy") == 0) {
            acc = (float)atof(v);
            hasAcc = true;
        } else if (strcmp(f, "--k") == 0) {
            K = (unsigned)atoi(v);
            if (K != 4 && K != 8 && K != 16 && K != 32) {
                std::cout << "K must be 4, 8, 16 or 32\n";
                exit(0);
            }
        } else if (strcmp(f, "--wgs") == 0) {
            wgsPerCU = (unsigned)atoi(v);
        } else if (strcmp(f, "--frontier") == 0) {
            if (strcmp(v, "sorted") == 0) frontierKind = EEG_FRONTIER_SORTED;
            else if (strcmp(v, "flags") == 0) frontierKind = EEG_FRONTIER_FLAGS;
            else if (strcmp(v, "compact") == 0) frontierKind = EEG_FRONTIER_COMPACT;
            else {
                std::cout << "Frontier must be sorted, compact or flags\n";
                exit(0);
            }
        } else if (strcmp(f, "--autotune") == 0) {
            if (is_true(v)) autotune = true;
        } else if (strcmp(f, "--algorithm") == 0) {
            algorithm = v;
        } else {
            std::cout << "\nThere was an error parsing command line argument <" << f << ">\n";
            std::cout << GenerateHelpString();
            exit(0);
        }
    }
    if (energy && (!hasEnergyFile || !hasEnergyStats)) {
        std::cout << "The option --energy was true but energy file and/or energy stats files were not included\n";
        exit(0);
    }
    if (!hasInput) {
        std::cout << "\nInput graph file argument is required.\n";
        std::cout << GenerateHelpString();
        exit(0);
    }
    return true;
}

std::string ArgumentParser::GenerateHelpString() {
    std::string str = "\nRequired arguments:";
    str += "\n    [--input]: Input graph file. E.g., --input FacebookGraph.txt";
    str += "\nOptional arguments";
    if (canHaveSource) str += "\n    [--source]:  Begins from the source (Default: 0). E.g., --source 10";
    str += "\n    [--output]: Output file for results. E.g., --output results.txt";
    str += "\n    [--device]: Select GPU device (Default: 0). E.g., --device 1";
    if (canHaveItrs) str += "\n    [--iteration]: Number of iterations (Default: 1). E.g., --iterations 10";
    str += "\n    [--debug]: Check or observe information (Default: false). E.g. --debug true";
    str += "\n    [--variant]: Algorithm variant option(Default: async_push_td). E.g. --variant async_push_td";
    str += "\n    [--algorithm]: Algorithm to run in `eegraph` (bfs, sssp, sswp, cc, pr). E.g. --algorithm bfs";
    str += "\n    [--energy]: Measure and output GPU energy information (Default: false). E.g. --energy true";
    str += "\n    [--efile]: Output file for energy (Required if energy == true). E.g. --efile my_experiment_energy";
    str += "\n    [--estats]: Output file for energy (Required if energy == true). E.g. --estats my_experiment_stats";
    str += "\n    [--um]: Accepted for compatibility; the graph always lives in HBM. E.g. --um true";
    str += "\n    [--accuracy]: PageRank L1 convergence threshold (Default: 1e-7). E.g. --accuracy 0.0000001";
    str += "\n    [--activity]: sum node activity vector in results(Default false) E.g. --activity true";
    str += "\n    [--k]: Virtual-node degree bound 4, 8, 16 or 32 (Default: 8). E.g. --k 16";
    str += "\n    [--wgs]: Workgroups per compute unit of the tile kernels (Default: auto). E.g. --wgs 16";
    str += "\n    [--frontier]: Frontier of sync_push_dd, sorted, compact or flags (Default: sorted)";
    str += "\n    [--autotune]: Pick K, workgroups and variant by timing short runs. E.g. --autotune true";
    str += "\n\n";
    return str;
}
