/*
 * host_check.cpp — runs the C++ host classes on inputs from a file, for the
 * tests that compare them with the oracle.
 *
 *   mpm_host_check mediator IN OUT
 *     IN:  L (3), cells (3), n, then n positions (3 each)       all doubles
 *     OUT: per position: count, 8 node indices (-1 past count), then per
 *          node N and the gradient (3)                          all doubles
 *   mpm_host_check constitutive IN OUT      (runs on the GPU)
 *     IN:  type, E, nu, sigma_y, C0, C1, n, then per point R, sigma (6), de (6)
 *     OUT: per point d_sigma (6), d_eps_p (6), dR
 */
#include <cstdio>
#include <cstring>
#include <vector>

#include "mpm_physics.h"

using namespace mpmhost;

static std::vector<double> read_all(const char *path) {
    std::vector<double> v;
    FILE *f = std::fopen(path, "rb");
    if (!f) return v;
    double x;
    while (std::fread(&x, 8, 1, f) == 1) v.push_back(x);
    std::fclose(f);
    return v;
}

int main(int argc, char **argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s mediator|constitutive IN OUT\n", argv[0]);
        return 2;
    }
    const std::vector<double> in = read_all(argv[2]);
    std::vector<double> out;
    if (!std::strcmp(argv[1], "mediator")) {
        if (in.size() < 7) return 2;
        GridPoint_Mediator gm;
        const dvec3 L(in[0], in[1], in[2]);
        const ivec3 cells((int)in[3], (int)in[4], (int)in[5]);
        gm.d3_Length_Cell = dvec3(L.x / cells.x, L.y / cells.y, L.z / cells.z);
        gm.i3_Node_Count = ivec3(cells.x + 1, cells.y + 1, cells.z + 1);
        GridPoint_Factory gf;
        std::vector<GridPoint *> grid = gf.createGrid(L, cells);
        const size_t n = (size_t)in[6];
        if (in.size() < 7 + 3 * n) return 2;
        for (size_t i = 0; i < n; i++) {
            const dvec3 x(in[7 + 3 * i], in[8 + 3 * i], in[9 + 3 * i]);
            gm.findAdjacentGridPoints(x);
            out.push_back((double)gm.v_adjacentGridPoints.size());
            for (size_t k = 0; k < 8; k++)
                out.push_back(k < gm.v_adjacentGridPoints.size() ? (double)gm.v_adjacentGridPoints[k] : -1.0);
            for (unsigned int node : gm.v_adjacentGridPoints) {
                gm.calculateBases_Classic(x, grid[node]->d3_Position);
                out.push_back(gm.d_ShapeValue);
                out.push_back(gm.d3_ShapeGradient.x);
                out.push_back(gm.d3_ShapeGradient.y);
                out.push_back(gm.d3_ShapeGradient.z);
            }
        }
        for (GridPoint *g : grid) delete g;
    } else if (!std::strcmp(argv[1], "constitutive")) {
        if (in.size() < 7) return 2;
        Material_BC m;
        m.i_MaterialType = (int)in[0];
        m.d_ElasticModulus = in[1];
        m.d_PoissonRatio = in[2];
        m.d_YieldStress = in[3];
        m.d_Hardening_Isotropic_C0 = in[4];
        m.d_Hardening_Isotropic_C1 = in[5];
        const size_t n = (size_t)in[6];
        if (in.size() < 7 + 13 * n) return 2;
        for (size_t i = 0; i < n; i++) {
            const double *r = &in[7 + 13 * i];
            double sig[6], de[6];
            for (int k = 0; k < 6; k++) {
                sig[k] = r[1 + k];
                de[k] = r[7 + k];
            }
            ConstitutiveRelation cr; /* a fresh object per point, as the step loops do */
            if (m.i_MaterialType == __ELASTIC)
                cr.calculateIncrement_Elastic(&m, de);
            else if (m.i_MaterialType == __VONMISESHARDENING)
                cr.calculateIncrement_VonMisesHardening(&m, r[0], sig, de);
            else
                cr.calculateIncrement_Plastic(&m, sig, de);
            for (int k = 0; k < 6; k++) out.push_back(cr.d6StressIncrement[k]);
            for (int k = 0; k < 6; k++) out.push_back(cr.d6PlasticStrainIncrement[k]);
            out.push_back(cr.dBackstress_IsotropicIncrement);
        }
    } else {
        return 2;
    }
    FILE *f = std::fopen(argv[3], "wb");
    if (!f) return 1;
    std::fwrite(out.data(), 8, out.size(), f);
    std::fclose(f);
    return 0;
}
