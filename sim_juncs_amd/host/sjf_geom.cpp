// sjf_geom.cpp -- command-line driver with the reference's sim_geom flags (src/main.cpp).
//
//   sjf_geom [--conf-file params.conf] [--geom-file junc.geom] [--out-dir DIR] [--grid-res R]
//            [--grid-num N] [--length L] [--eps1 E] [--save-span S] [-v V] [--opts "a=1"]
//            [--fp32] [--real] [--device D]
//
// The geometry file is a .geom scene script (run with the settings and --opts, as the
// reference does) or a scene JSON (roots, sources, monitor clusters, numeric context entries;
// file name ending in .json).  Steps, as main():
// parse_args, parse_conf_file (params.conf unless --conf-file), correct_defaults, build the
// simulation, run, save field_samples.h5, printing the same timing lines.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "sjf_host.hpp"

int main(int argc, char** argv) {
    using namespace sjfh;
    RunOptions opt;
    std::vector<char*> rest;
    for (int i = 0; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--fp32")) opt.precision = SJF_FP32;
        else if (!std::strcmp(argv[i], "--real")) opt.complex_fields = false;
        else if (!std::strcmp(argv[i], "--device") && i + 1 < argc) opt.device = std::atoi(argv[++i]);
        else rest.push_back(argv[i]);
    }
    int n = (int)rest.size();
    Settings args;
    int ret = parse_args(&args, &n, rest.data());
    if (ret) return ret;
    const std::string conf = args.has_conf_fname ? args.conf_fname : std::string("params.conf");
    ret = parse_conf_file(&args, conf);
    if (ret) {
        std::fprintf(stderr, "cannot read %s\n", conf.c_str());
        return 1;
    }
    correct_defaults(&args);
    try {
        const auto start = std::chrono::steady_clock::now();
        const Scene scene = load_scene(args.geom_fname, args);
        BoundGeom geom(args, scene, opt);
        const auto end_init = std::chrono::steady_clock::now();
        std::cout << "initialization completed in "
                  << std::chrono::duration_cast<std::chrono::milliseconds>(end_init - start).count() << " ms" << std::endl;
        geom.run(args.out_dir);
        const auto end_run = std::chrono::steady_clock::now();
        std::cout << "simulation completed in " << std::chrono::duration_cast<std::chrono::seconds>(end_run - end_init).count()
                  << " s" << std::endl;
        geom.save_field_times(args.out_dir);
        const auto end_write = std::chrono::steady_clock::now();
        std::cout << "saving timeseries completed in "
                  << std::chrono::duration_cast<std::chrono::milliseconds>(end_write - end_run).count() << " ms" << std::endl;
        std::cout << "total time: " << std::chrono::duration_cast<std::chrono::seconds>(end_write - start).count() << " s"
                  << std::endl;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
