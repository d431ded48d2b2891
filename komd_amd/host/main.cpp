// komd_md: runs a KoMD deck on the GPU path.
//
//   komd_md DECK [--steps N] [--device-loop] [--check] [--write PATH]
//
//   --check        parse the deck and its checkpoint files and print a summary; no GPU use
//   --write PATH   write the sites as a .ps checkpoint (after the run, or as loaded with --check)
//   --device-loop  run the steps as one device-queued loop (komd_run) instead of the
//                  reference's per-functor loop
// Prints one JSON line: sites, molecules, grid, steps, temperature and energies.
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>

#include "komd_host.hpp"

using namespace komd_host;

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s DECK [--steps N] [--device-loop] [--check] [--write PATH]\n", argv[0]);
        return 2;
    }
    std::string deck = argv[1], write_path;
    long long steps = -1;
    bool check_only = false, device_loop = false;
    for (int i = 2; i < argc; i++) {
        if (!std::strcmp(argv[i], "--steps") && i + 1 < argc) steps = std::atoll(argv[++i]);
        else if (!std::strcmp(argv[i], "--check")) check_only = true;
        else if (!std::strcmp(argv[i], "--device-loop")) device_loop = true;
        else if (!std::strcmp(argv[i], "--write") && i + 1 < argc) write_path = argv[++i];
        else {
            std::fprintf(stderr, "unknown argument %s\n", argv[i]);
            return 2;
        }
    }
    try {
        Configuration config;
        std::vector<Component> comps;
        std::string err;
        if (!readFile(deck, config, comps, &err)) {
            std::fprintf(stderr, "%s: %s\n", deck.c_str(), err.c_str());
            return 1;
        }
        if (steps < 0) steps = (long long)config.timesteps;
        // the pair functor the registry puts in LJ's slot (main.cpp:56-57)
        const char* pair = !config.pair_table.empty() ? "table" : (config.enable_3b_approx ? "atm2b" : "lj12_6");
        if (check_only) {
            Sites sites;
            CheckpointIO::loadCheckpoint(config, sites);
            uint64_t mols = 0;
            for (size_t i = 0; i < sites.size(); i++)
                if (i == 0 || sites.id[i] != sites.id[i - 1]) mols++;
            if (!write_path.empty()) CheckpointIO::writePs(write_path, sites, config.checkpoint_precision);
            std::printf(
                "{\"sites\": %zu, \"molecules\": %llu, \"components\": %zu, \"cutoff\": %.17g, \"cell_size\": %.17g, "
                "\"delta_t\": %.17g, \"temperature\": %.17g, \"nu_internal\": %.17g, \"limit\": %.17g, "
                "\"domain_high\": [%.17g, %.17g, %.17g], \"enable_3b\": %d, \"atm_mode\": \"%s\", \"timesteps\": %lld, "
                "\"first_site\": [%.17g, %.17g, %.17g], \"pair\": \"%s\", \"pair_steps\": %d, \"triple_steps\": %d, "
                "\"skin\": %.17g}\n",
                sites.size(), (unsigned long long)mols, comps.size(), config.cutoff, config.cell_size, config.delta_t,
                config.temperature, config.energy_3b * 1e9 * kConvJEi, config.limit_factor, config.domainHigh[0],
                config.domainHigh[1], config.domainHigh[2], (int)config.enable_3b, config.atm_mode.c_str(), steps,
                sites.size() ? sites.r[0] : 0.0, sites.size() ? sites.r[1] : 0.0, sites.size() ? sites.r[2] : 0.0, pair,
                config.pair_steps, config.triple_steps, config.skin);
            return 0;
        }
        auto reg = makeRegistry(config);
        Simulation sim;
        if (device_loop) sim.runDevice(*reg, (uint64_t)steps);
        else sim.run(*reg, (uint64_t)steps);
        komd_observables o{};
        int kind = 0;
        if (reg->container->threeBody() && !reg->forceFunctors3b.empty()) {
            if (auto a = std::dynamic_pointer_cast<ATM>(reg->forceFunctors3b[0])) kind = a->kind();
            if (auto s = std::dynamic_pointer_cast<ATMSurrogate>(reg->forceFunctors3b[0])) kind = s->kind();
        }
        check(reg->container->ctx(), komd_measure_observables(reg->container->ctx(), kind, &o), "observables");
        auto dims = static_cast<LinkedCells&>(*reg->container).gridDims();
        if (!write_path.empty() || config.storeCheckpoint)
            CheckpointIO::writeCheckpoint(sim.simstep(), *reg->container, config.checkpoint_precision, write_path);
        std::printf(
            "{\"sites\": %zu, \"molecules\": %llu, \"grid\": [%lld, %lld, %lld], \"steps\": %llu, \"loop\": \"%s\", "
            "\"temperature\": %.17g, \"kinetic\": %.17g, \"u_lj\": %.17g, \"u_3b\": %.17g, \"virial\": %.17g, "
            "\"pairs\": %lld, \"triplets\": %lld, \"sensors\": %zu, \"pair\": \"%s\"}\n",
            reg->container->sites().size(), (unsigned long long)reg->container->getNumMolecules(), (long long)dims[0],
            (long long)dims[1], (long long)dims[2], (unsigned long long)sim.simstep(),
            device_loop ? "device" : "reference", o.temperature, o.kinetic, o.u_lj, o.u_3b, o.virial,
            (long long)o.pairs, (long long)o.triplets, reg->sensors.size(), pair);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "komd_md: %s\n", e.what());
        return 1;
    }
    return 0;
}
