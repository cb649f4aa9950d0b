// run_model.cpp -- C++ host driver over include/molflow_gpu.hpp, used by the
// GPU tests to exercise the boundary from C++ the way a C++ simulation
// manager would: check and load a model, build the BVH, reset, run steps
// (optionally keeping particles in flight between them) with a periodic
// report after each (UpdateHitsAndLog: add into the host state, clear the
// device tallies), drain, and report once more.
//
//   run_model MODEL_FILE OUT_FILE SEED N_STEPS DES_PER_STEP CARRY(0|1)
//
// MODEL_FILE: "MFGM1\0\0\0", u32 n_vertices, n_facets, n_index, 0; then
// f64 vertices[3*n_vertices], u32 offsets[n_facets+1], u32 index[n_index],
// mfg_facet_desc[n_facets], mfg_settings (raw structs of molflow_gpu.h).
// OUT_FILE (text): "hits H", "desorbed D", "abs_equiv A", "in_flight_max M",
// then one line "f nb_mc_hit nb_abs_equiv" per facet.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "molflow_gpu.hpp"

namespace {
template <typename T> void read_into(std::ifstream& in, T* p, size_t n) {
    in.read(reinterpret_cast<char*>(p), (std::streamsize)(n * sizeof(T)));
    if (!in) throw std::runtime_error("model file truncated");
}
}  // namespace

int main(int argc, char** argv) {
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s MODEL OUT SEED N_STEPS DES_PER_STEP CARRY\n", argv[0]);
        return 2;
    }
    try {
        std::ifstream in(argv[1], std::ios::binary);
        char magic[8];
        read_into(in, magic, 8);
        if (std::memcmp(magic, "MFGM1\0\0\0", 8) != 0) throw std::runtime_error("not a model file");
        uint32_t hdr[4];
        read_into(in, hdr, 4);
        std::vector<double> verts(3 * (size_t)hdr[0]);
        std::vector<uint32_t> off(hdr[1] + 1), idx(hdr[2]);
        std::vector<mfg_facet_desc> facets(hdr[1]);
        mfg_settings settings{};
        read_into(in, verts.data(), verts.size());
        read_into(in, off.data(), off.size());
        read_into(in, idx.data(), idx.size());
        read_into(in, facets.data(), facets.size());
        read_into(in, &settings, 1);

        mfg_model_view v{};
        v.struct_size = sizeof(v);
        v.abi_version = MFG_ABI_VERSION;
        v.n_vertices = hdr[0];
        v.n_facets = hdr[1];
        v.n_structures = 1;
        v.vertices = verts.data();
        v.facet_vertex_offset = off.data();
        v.facet_vertex_index = idx.data();
        v.facets = facets.data();
        v.settings = &settings;

        const uint64_t seed = std::strtoull(argv[3], nullptr, 10);
        const uint64_t steps = std::strtoull(argv[4], nullptr, 10);
        const uint64_t per = std::strtoull(argv[5], nullptr, 10);
        const bool carry = std::atoi(argv[6]) != 0;

        const std::vector<std::string> problems = mfg::Simulation::SanityCheckModel(v);
        if (!problems.empty()) throw std::runtime_error("model check: " + problems.front());
        mfg::Simulation sim(0);
        sim.LoadSimulation(v);
        sim.RebuildAccelStructure();
        sim.ResetSimulation(seed);
        sim.SetCarry(carry);
        // the host's global state (GlobalSimuState role)
        mfg_tally_layout L = sim.Layout();
        mfg_global_counters g{};
        std::vector<mfg_facet_counters> fc((size_t)L.n_slots * L.n_facets);
        std::memset(fc.data(), 0, fc.size() * sizeof(mfg_facet_counters));
        mfg_tally_view tv{};
        tv.struct_size = sizeof(tv);
        tv.global = &g;
        tv.facets = fc.data();
        uint64_t remaining = steps * per, in_flight_max = 0;
        for (;;) {
            mfg::MCStepResult r = sim.SimulationMCStep(per, remaining, 1000000);
            if (r == mfg::MCStepResult::DesorptionError) throw std::runtime_error("desorption error");
            remaining -= per < remaining ? per : remaining;
            uint64_t fl = sim.InFlight();
            if (fl > in_flight_max) in_flight_max = fl;
            sim.UpdateHitsAndLog(tv);           // periodic report
            if (r == mfg::MCStepResult::MaxReached) break;
        }
        sim.Drain(1000000);
        sim.UpdateHitsAndLog(tv);

        std::FILE* out = std::fopen(argv[2], "w");
        if (!out) throw std::runtime_error("cannot write output");
        std::fprintf(out, "hits %llu\ndesorbed %llu\nabs_equiv %.17g\nin_flight_max %llu\n",
                     (unsigned long long)g.nb_mc_hit, (unsigned long long)g.nb_desorbed, g.nb_abs_equiv,
                     (unsigned long long)in_flight_max);
        for (uint32_t f = 0; f < L.n_facets; ++f)
            std::fprintf(out, "%u %llu %.17g\n", f, (unsigned long long)fc[f].nb_mc_hit, fc[f].nb_abs_equiv);
        std::fclose(out);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "run_model: %s\n", e.what());
        return 1;
    }
    return 0;
}
