// Executables bfs, sssp, sswp, cc, pr (one build of this file each, EEG_MAIN_ALGO set by the
// Makefile) and `eegraph` (algorithm from --algorithm).  Follows src/{bfs,sssp,sswp,cc,pr}_main.cu:
// parse arguments, read the graph, run the algorithm.  Exit status 0 on success (including the
// reference's help/argument exits), 1 when the graph or the run fails.
#include <cstring>
#include <iostream>

#include "eegraph_host.hpp"

#ifndef EEG_MAIN_ALGO
#define EEG_MAIN_ALGO -1
#endif

int main(int argc, char** argv) {
    ArgumentParser arguments(argc, argv, true, false);
    int algo = EEG_MAIN_ALGO;
    if (algo < 0) {
        const std::string& a = arguments.algorithm;
        if (a == "bfs") algo = EEG_BFS;
        else if (a == "sssp" || a.empty()) algo = EEG_SSSP;  // the help text's default
        else if (a == "sswp") algo = EEG_SSWP;
        else if (a == "cc") algo = EEG_CC;
        else if (a == "pr") algo = EEG_PR;
        else {
            std::cout << "Algorithm not recognized\n";
            return 0;
        }
    }
    // the reference reads weighted input for every algorithm (src/sssp_main.cu:263)
    Graph graph(arguments.input, true);
    if (!graph.ReadGraph()) return 1;
    if (graph.num_nodes < 1) {
        std::cout << "Graph file not read correctly" << std::endl;
        return 1;
    }
    Result r = eegraph_run((uint32_t)algo, arguments, graph);
    if (arguments.nodeActivity && r.status == EEG_OK) {  // as src/main.cu:171-174 prints them
        std::cout << "Max active pct: " << utilities::maxActivePct(r.sumLabelsVec, graph.num_nodes) << std::endl;
        std::cout << "Pct over threshold: " << utilities::pctIterOverThreshold(r.sumLabelsVec, graph.num_nodes, .50)
                  << std::endl;
    }
    return (r.status == EEG_OK) ? 0 : 1;
}
