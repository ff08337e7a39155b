/*
 * bar_demo.cpp — a reference-style driver over the C++ host classes: builds
 * the cantilever bar with PhysicsEngine::initializeWorld_Classic_Bar, runs
 * runSimulation_Classic_DoublePass_MPLocks for the requested number of
 * steps' worth of time, and writes the particles' state (and the console
 * reports) to a binary file.
 *
 *   mpm_bar_demo STEPS OUTFILE [ORDER [PHASES]]
 *       ORDER: 0 particle, 1 cell (default 0); STEPS 0: the world as built,
 *       no device use; PHASES 1: each step driven phase by phase
 *       (runSimulation_Phase) instead of by the loop (no console reports
 *       besides the scene's own at t = 0)
 *
 * File: int64 n, int64 reports; then n-major doubles per field, component
 * by component (x 3, v 3, sigma 6, F 9, strain energy 1, volume 1); then 9
 * doubles per console report.
 */
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mpm_physics.h"

using namespace mpmhost;

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s STEPS OUTFILE [ORDER [PHASES]]\n", argv[0]);
        return 2;
    }
    const int steps = std::atoi(argv[1]);
    PhysicsEngine pe(0);
    pe.initializeWorld_Classic_Bar();
    pe.i_SummationOrder = argc > 3 ? std::atoi(argv[3]) : MPM_ORDER_PARTICLE;
    const bool phases = argc > 4 && std::atoi(argv[4]) != 0;
    /* STEPS = 0 writes the world as built, without touching the device */
    int status = 0;
    if (steps > 0 && !phases) status = pe.runSimulation_Classic_DoublePass_MPLocks(steps * pe.getTime_Increment());
    for (int s = 0; phases && s < steps && status >= 0; s++) {
        static const int32_t order[] = {MPM_PHASE_BIN, MPM_PHASE_P2G_GRID, MPM_PHASE_G2P_VELOCITY,
                                        MPM_PHASE_REMAP_GRID, MPM_PHASE_G2P_STRAIN};
        for (int32_t ph : order)
            if (status >= 0) status = pe.runSimulation_Phase(MPM_VARIANT_DOUBLEPASS, ph, pe.getTime_Increment());
    }
    if (status < 0) {
        std::fprintf(stderr, "runSimulation failed: %s\n", pe.last_error());
        return 1;
    }
    const std::vector<MaterialPoint_BC *> mp = pe.getMaterialPoints();
    const int64_t n = (int64_t)mp.size(), nr = (int64_t)pe.v_ConsoleLog.size();
    FILE *f = std::fopen(argv[2], "wb");
    if (!f) return 1;
    std::fwrite(&n, 8, 1, f);
    std::fwrite(&nr, 8, 1, f);
    std::vector<double> col((size_t)n);
    auto put = [&](auto get) {
        for (int64_t i = 0; i < n; i++) col[(size_t)i] = get(mp[(size_t)i]);
        std::fwrite(col.data(), 8, (size_t)n, f);
    };
    put([](MaterialPoint_BC *p) { return p->d3_Position.x; });
    put([](MaterialPoint_BC *p) { return p->d3_Position.y; });
    put([](MaterialPoint_BC *p) { return p->d3_Position.z; });
    put([](MaterialPoint_BC *p) { return p->d3_Velocity.x; });
    put([](MaterialPoint_BC *p) { return p->d3_Velocity.y; });
    put([](MaterialPoint_BC *p) { return p->d3_Velocity.z; });
    for (int k = 0; k < 6; k++) put([k](MaterialPoint_BC *p) { return p->d6_Stress[k]; });
    for (int k = 0; k < 9; k++) put([k](MaterialPoint_BC *p) { return p->d33_DeformationGradient.m[k / 3][k % 3]; });
    put([](MaterialPoint_BC *p) { return p->d_Energy_Strain; });
    put([](MaterialPoint_BC *p) { return p->d_Volume; });
    for (const ConsoleReport &r : pe.v_ConsoleLog) {
        const double row[9] = {r.time, r.position_y, r.force_y, r.momentum_y, r.kinetic_energy,
                               r.strain_energy, r.plastic_energy, r.strain_y, r.stress_y};
        std::fwrite(row, 8, 9, f);
    }
    std::fclose(f);
    std::printf("bar: %lld points, status %d, t = %.9e, %lld console reports\n", (long long)n, status,
                pe.getTime_Current(), (long long)nr);
    return 0;
}
