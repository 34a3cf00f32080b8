// Command line of a GSWITCH application.
//
// Same options and globals as the reference (cmd_t / cmd_opt / parse_cmd,
// src/utils/cmdline.cuh:38-232; LOG, src/utils/utils.cuh:21-23; mwtime,
// utils.cuh:176), so an application's main() keeps its shape:
//   parse_cmd(argc, argv, "BFS");
//   auto g = build_graph<VC>(cmd_opt.path, fets, cmd_opt.with_header,
//                            cmd_opt.with_weight, cmd_opt.directed);
// `--configs=Dir-Fmt-LB-Fusion` pins the selector's choices: init_conf copies
// cmd_opt.ins into config_t::force, where every selector reads it.
#pragma once
#include <getopt.h>
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "../abstraction/config.h"
#include "common.h"

struct cmd_t {
  int src = -1;
  std::string configs = "";
  std::string path = "";
  std::string json = "";
  bool verbose = false;
  bool validation = false;
  bool with_header = false;  // first non-comment line is "rows cols nnz"
  bool with_weight = false;  // third column holds an integer weight
  bool directed = false;
  int device = 0;
  ins_t ins;

  void output() const {
    std::cout << " src: " << src << "\n verbose: " << (verbose ? "True" : "False")
              << "\n validation: " << (validation ? "True" : "False")
              << "\n with_header: " << (with_header ? "True" : "False")
              << "\n with_weight: " << (with_weight ? "True" : "False")
              << "\n directed: " << (directed ? "True" : "False")
              << "\n configs: " << configs << "\n path: " << path
              << "\n device: " << device << std::endl;
  }
};

inline cmd_t cmd_opt;

// wall-clock milliseconds
inline double mwtime() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch())
      .count();
}

#define LOG(...)                                   \
  do {                                             \
    if (cmd_opt.verbose) std::printf(__VA_ARGS__); \
  } while (0)

inline void cmd_help(const std::string &exec) {
  std::cout << "./" << exec << " <graph_path> [options]\n"
            << "  -r, --src=<int>       root vertex (default: a random vertex "
               "with edges)\n"
            << "  -v, --verbose         per-level log\n"
            << "  -V, --validation      check the result on the CPU\n"
            << "  -H, --with-header     the file has a size line\n"
            << "  -W, --with-weight     the file has edge weights\n"
            << "  -d, --directed        the graph is directed\n"
            << "  -c, --configs=Push-Queue-CM-Fused  pin strategies "
               "(x = free)\n"
            << "  -D, --device=<int>    GPU to run on\n"
            << "  -j, --json=<path>     append per-level training records"
            << std::endl;
}

// argv[1] is the graph file; options follow it. Prints the help and exits
// on a missing file or an unknown option, as the reference does.
inline void parse_cmd(int argc, char **argv, const std::string &exec = "test") {
  if (argc < 2) {
    cmd_help(exec);
    std::exit(0);
  }
  cmd_opt.path = argv[1];
  struct stat st;
  if (::stat(cmd_opt.path.c_str(), &st) != 0) {
    std::cout << " file \"" << cmd_opt.path << "\" does not exist." << std::endl;
    std::exit(0);
  }
  static const struct option longopts[] = {
      {"src", required_argument, 0, 'r'},  {"verbose", no_argument, 0, 'v'},
      {"validation", no_argument, 0, 'V'}, {"with-header", no_argument, 0, 'H'},
      {"with-weight", no_argument, 0, 'W'}, {"directed", no_argument, 0, 'd'},
      {"configs", required_argument, 0, 'c'}, {"json", required_argument, 0, 'j'},
      {"device", required_argument, 0, 'D'}, {0, 0, 0, 0},
  };
  optind = 1;  // getopt restarts (parse_cmd may run more than once)
  int idx = 0, opt;
  while ((opt = getopt_long(argc - 1, &argv[1], "r:vVHWdc:j:D:", longopts,
                            &idx)) != -1) {
    switch (opt) {
      case 'r': cmd_opt.src = std::atoi(optarg); break;
      case 'v': cmd_opt.verbose = true; break;
      case 'V': cmd_opt.validation = true; break;
      case 'H': cmd_opt.with_header = true; break;
      case 'W': cmd_opt.with_weight = true; break;
      case 'd': cmd_opt.directed = true; break;
      case 'c': cmd_opt.configs = optarg; break;
      case 'j': cmd_opt.json = optarg; break;
      case 'D': cmd_opt.device = std::atoi(optarg); break;
      default:
        cmd_help(exec);
        std::exit(0);
    }
  }
  cmd_opt.ins = parse_configs(cmd_opt.configs);
  H_ERR(hipSetDevice(cmd_opt.device));
}

// The forced choices of the command line, if any (init_conf applies them).
inline bool cmd_forces_anything(const ins_t &i) {
  return i.has_direction || i.has_asfmt || i.has_lb || i.has_fusion || i.has_sorted;
}
