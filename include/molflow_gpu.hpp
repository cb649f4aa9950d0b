/*
 * molflow_gpu.hpp -- C++ host shim over the C ABI of molflow_gpu.h.
 *
 * mfg::Simulation plays the role of MolflowSimulation : Simulation_Abstract
 * (reference src/Simulation/MolflowSimulation.h:22-45) for a host that is
 * itself C++: the same method names, one GPU simulation unit per object.
 * Errors become mfg::Error exceptions on this side of the ABI; nothing
 * throws across it. Header-only; link with -lmolflow_gpu.
 */
#ifndef MOLFLOW_GPU_HPP
#define MOLFLOW_GPU_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "molflow_gpu.h"

namespace mfg {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }
private:
    int code_;
};

/* reference src/Simulation/Particle.h:24-28 */
enum class MCStepResult { Success, DesorptionError, MaxReached };

class Simulation {
public:
    /* ConstructParticleTracers (MolflowSimulation.h:40): one GPU unit */
    explicit Simulation(int device = 0) {
        int rc = mfg_create(&ctx_, device);
        if (rc != MFG_OK) throw Error(rc, "mfg_create: no usable HIP device " + std::to_string(device));
    }
    ~Simulation() { if (ctx_) mfg_destroy(ctx_); }
    Simulation(const Simulation&) = delete;
    Simulation& operator=(const Simulation&) = delete;

    /* LoadSimulation + PrepareToRun (MolflowSimulation.h:29, MolflowSimGeom.cpp:194-335) */
    void LoadSimulation(const mfg_model_view& view) { check(mfg_load_model(ctx_, &view), "mfg_load_model"); }

    /* RebuildAccelStructure (MolflowSimulation.h:30) */
    int RebuildAccelStructure(int width = 8, int leafSize = 2) {
        check(mfg_build_accel(ctx_, width, leafSize), "mfg_build_accel");
        return 0;
    }

    /* ResetSimulation (MolflowSimulation.h:32) */
    void ResetSimulation(uint64_t seed, uint64_t particleOffset = 0) {
        check(mfg_reset(ctx_, seed, particleOffset), "mfg_reset");
    }

    /* GetHitsSize (MolflowSimulation.h:34) */
    size_t GetHitsSize() const {
        uint64_t b = 0;
        check(mfg_hits_size(ctx_, &b), "mfg_hits_size");
        return (size_t)b;
    }

    /* SimulationMCStep (Particle.h:45): desorbs up to min(nbDesorptions,
     * remainingDes) new particles. With carry enabled, particles still in
     * flight stay on the device for the next call, like the reference's
     * tracer state between calls (Particle.h:116-117). */
    MCStepResult SimulationMCStep(uint64_t nbDesorptions, uint64_t remainingDes, uint64_t maxEvents = 0,
                                  mfg_step_result* out = nullptr) {
        uint64_t n = nbDesorptions < remainingDes ? nbDesorptions : remainingDes;
        if (n == 0) return MCStepResult::MaxReached;
        mfg_step_result r{};
        int rc = mfg_run(ctx_, n, maxEvents, &r);
        if (rc == MFG_ERR_DESORPTION) return MCStepResult::DesorptionError;
        check(rc, "mfg_run");
        if (out) *out = r;
        return n == remainingDes ? MCStepResult::MaxReached : MCStepResult::Success;
    }

    /* exitRequested (Particle.h:122) */
    void RequestStop() { check(mfg_request_stop(ctx_), "mfg_request_stop"); }
    /* Wall-clock limit of one trace launch (a run it ends reports stopped = 2) */
    void SetTimeLimit(double seconds) { check(mfg_set_time_limit(ctx_, seconds), "mfg_set_time_limit"); }

    /* UpdateMCHits (Particle.cpp:21-73): adds the device tallies into the
     * caller's buffers (laid out as GlobalSimuState, see molflow_gpu.h) */
    void UpdateMCHits(const mfg_tally_view& view) const { check(mfg_read_tallies(ctx_, &view), "mfg_read_tallies"); }

    /* UpdateHitsAndLog (Particle.cpp:1294-1312): adds the device tallies
     * into the caller's buffers, then clears them on the device (the
     * tmpState->Reset() step); particles in flight and the particle stream
     * are kept, so periodic reports add up to one uninterrupted run */
    void UpdateHitsAndLog(const mfg_tally_view& view) { check(mfg_take_tallies(ctx_, &view), "mfg_take_tallies"); }

    /* ReinitializeParticleLog (MolflowSimulation.h:35) */
    int ReinitializeParticleLog() {
        check(mfg_reinit_particle_log(ctx_), "mfg_reinit_particle_log");
        return 0;
    }

    /* SanityCheckModel (MolflowSimulation.h:29): problems that stop the
     * model from loading plus the reference's own warnings; no device
     * needed */
    static std::vector<std::string> SanityCheckModel(const mfg_model_view& view) {
        std::vector<char> buf(1 << 16);
        uint32_t n = 0;
        if (mfg_sanity_check(&view, buf.data(), buf.size(), &n) != MFG_OK) return {"mfg_sanity_check failed"};
        std::vector<std::string> out;
        std::string cur;
        for (const char* p = buf.data(); *p; ++p) {
            if (*p == '\n') { out.push_back(cur); cur.clear(); }
            else cur += *p;
        }
        if (!cur.empty()) out.push_back(cur);
        return out;
    }

    mfg_tally_layout Layout() const {
        mfg_tally_layout L{};
        L.struct_size = sizeof(L);
        check(mfg_get_layout(ctx_, &L), "mfg_get_layout");
        return L;
    }

    void SetCarry(bool on) { check(mfg_set_carry(ctx_, on ? 1 : 0), "mfg_set_carry"); }
    uint64_t InFlight() const {
        uint64_t n = 0;
        check(mfg_in_flight(ctx_, &n), "mfg_in_flight");
        return n;
    }
    /* traces every parked particle to its end */
    mfg_step_result Drain(uint64_t maxEvents = 0) {
        SetCarry(false);
        mfg_step_result r{};
        check(mfg_run(ctx_, 0, maxEvents, &r), "mfg_run");
        return r;
    }
    mfg_tune_result Autotune(uint64_t nProbe) {
        mfg_tune_result t{};
        check(mfg_autotune(ctx_, nProbe, &t), "mfg_autotune");
        return t;
    }

    mfg_context* handle() const { return ctx_; }

private:
    void check(int rc, const char* what) const {
        if (rc != MFG_OK) {
            const char* m = mfg_last_error(ctx_);
            throw Error(rc, std::string(what) + ": " + (m ? m : ""));
        }
    }
    mfg_context* ctx_ = nullptr;
};

}  // namespace mfg

#endif
