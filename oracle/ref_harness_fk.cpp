// ref_harness_fk.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Drives the reference's own FK operator and update functor, compiled from
// the reference sources where they lie (Glioma/Glioma_ReactionDiffusionOperator.h,
// MRAGcore/MRAGCommon.h), on caller-supplied ghosted labs.  Built twice, with
// _MRAG_REAL_=float (the reference default) and _MRAG_REAL_=double, into
// oracle/_ref/.  Used by tests to pin oracle/fk_oracle.c bit for bit.
//
// The lab and block types below are the template arguments the reference's
// functor concept expects (MRAGmultithreading/MRAGBlockProcessing_SingleCPU.h:
// 92-155): operator()(ix,iy,iz) returning an element with phi, dphidt, p_w, p_g.
#include "MRAGcore/MRAGCommon.h"
#undef max
#undef min
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif
using namespace std;
using namespace MRAG;
#include "Glioma/Glioma_ReactionDiffusionOperator.h"

#define GQ_CAT2(a, b) a##_##b
#define GQ_CAT(a, b) GQ_CAT2(a, b)
#define GQ_EXPORT extern "C" __attribute__((visibility("default")))

namespace {
struct HCell {
    Real phi, dphidt, p_g, p_w;
};

template <int B>
struct HLab {
    HCell *d;
    HCell &operator()(int ix, int iy = 0, int iz = 0) { return d[(ix + 1) + (B + 2) * ((iy + 1) + (B + 2) * (iz + 1))]; }
};

template <int B>
struct HBlock {
    static const int sizeX = B, sizeY = B, sizeZ = B;
    HCell *d;
    HCell &operator()(int ix, int iy = 0, int iz = 0) { return d[ix + B * (iy + B * iz)]; }
};

template <int B>
void run_fk(const Real *lp, const Real *lw, const Real *lg, Real h, Real Dw, Real Dg, Real rho, Real *dphidt)
{
    const int L = B + 2, n = L * L * L;
    HCell *lab = new HCell[n];
    HCell *blk = new HCell[B * B * B];
    for (int i = 0; i < n; i++) { lab[i].phi = lp[i]; lab[i].p_w = lw[i]; lab[i].p_g = lg[i]; lab[i].dphidt = 0; }
    for (int i = 0; i < B * B * B; i++) { blk[i].phi = 0; blk[i].p_w = 0; blk[i].p_g = 0; blk[i].dphidt = 0; }
    HLab<B> l = {lab};
    HBlock<B> o = {blk};
    const int idx[3] = {0, 0, 0};
    const Real pos[3] = {0, 0, 0};
    const Real sp[3] = {h, h, h};
    BlockInfo info(0, idx, 0, pos, sp);
    Glioma_ReactionDiffusionOperator<3> rhs(Dw, Dg, rho);
    rhs(l, info, o);
    for (int i = 0; i < B * B * B; i++) dphidt[i] = blk[i].dphidt;
    delete[] lab;
    delete[] blk;
}

template <int B>
void run_update(Real *phi, const Real *dphidt, double dt)
{
    HCell *blk = new HCell[B * B * B];
    for (int i = 0; i < B * B * B; i++) { blk[i].phi = phi[i]; blk[i].dphidt = dphidt[i]; blk[i].p_w = 0; blk[i].p_g = 0; }
    HBlock<B> o = {blk};
    const int idx[3] = {0, 0, 0};
    BlockInfo info(0, idx, 0);
    UpdateTumor<3> up(dt);
    up(info, o);
    for (int i = 0; i < B * B * B; i++) phi[i] = blk[i].phi;
    delete[] blk;
}
}  // namespace

GQ_EXPORT int GQ_CAT(gqref_fk_block, GQREF_SUFFIX)(const Real *lab_phi, const Real *lab_pw, const Real *lab_pg, int B,
                                                  Real h, Real Dw, Real Dg, Real rho, Real *dphidt)
{
    switch (B) {
    case 4: run_fk<4>(lab_phi, lab_pw, lab_pg, h, Dw, Dg, rho, dphidt); return 0;
    case 8: run_fk<8>(lab_phi, lab_pw, lab_pg, h, Dw, Dg, rho, dphidt); return 0;
    case 16: run_fk<16>(lab_phi, lab_pw, lab_pg, h, Dw, Dg, rho, dphidt); return 0;
    default: return -1;
    }
}

GQ_EXPORT int GQ_CAT(gqref_update, GQREF_SUFFIX)(Real *phi, const Real *dphidt, int B, double dt)
{
    switch (B) {
    case 4: run_update<4>(phi, dphidt, dt); return 0;
    case 8: run_update<8>(phi, dphidt, dt); return 0;
    case 16: run_update<16>(phi, dphidt, dt); return 0;
    default: return -1;
    }
}

GQ_EXPORT int GQ_CAT(gqref_real_size, GQREF_SUFFIX)() { return (int)sizeof(Real); }

// ---- CPU baseline driver -------------------------------------------------
// n_steps of _reactionDiffusionStep (Glioma/Glioma_HG_UQ.cpp:266-277) on a
// uniform periodic grid of N^3 points in 16^3 blocks, with the reference's
// own functors: per block a lab of the RD stencil [-1,2) is gathered from the
// block and its same-level neighbours (what BlockLab::load does on a uniform
// grid, MRAGcore/MRAGBlockLab.h:154-237), the RD functor writes dphidt into the
// block, then UpdateTumor updates every block (BlockProcessing::process).
// Cells have the 41-Real footprint of CCell_extended (Glioma/Glioma_Types.h:
// 22-108), so the memory traffic is the reference's; OpenMP over blocks
// stands in for the TBB pipeline.  Volumes are [z][y][x]; phi is updated in
// place.
namespace {
struct Cell41 {
    Real phi, dphidt, p_g, p_w;
    Real other[37];
};

struct LabD {
    Cell41 *d;
    Cell41 &operator()(int ix, int iy = 0, int iz = 0) { return d[(ix + 1) + 18 * ((iy + 1) + 18 * (iz + 1))]; }
};

struct BlockD {
    static const int sizeX = 16, sizeY = 16, sizeZ = 16;
    Cell41 *d;
    Cell41 &operator()(int ix, int iy = 0, int iz = 0) { return d[ix + 16 * (iy + 16 * iz)]; }
};
}  // namespace

GQ_EXPORT int GQ_CAT(gqref_uniform_steps, GQREF_SUFFIX)(Real *phi, const Real *pw, const Real *pg, int N, double h,
                                                        Real Dw, Real Dg, Real rho, double dt, int n_steps, int threads)
{
    if (N % 16) return -1;
    const int nb = N / 16, NB = nb * nb * nb;
    std::vector<Cell41> grid((size_t)NB * 4096);
    for (int b = 0; b < NB; b++) {
        const int bx = b % nb, by = (b / nb) % nb, bz = b / (nb * nb);
        for (int i = 0; i < 4096; i++) {
            const size_t v = (size_t)(bx * 16 + i % 16) + N * ((size_t)(by * 16 + (i / 16) % 16) + (size_t)N * (bz * 16 + i / 256));
            Cell41 &c = grid[(size_t)b * 4096 + i];
            std::memset(&c, 0, sizeof(c));
            c.phi = phi[v];
            c.p_w = pw[v];
            c.p_g = pg[v];
        }
    }
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    const Glioma_ReactionDiffusionOperator<3> rhs(Dw, Dg, rho);
    const UpdateTumor<3> up(dt);
    for (int s = 0; s < n_steps; s++) {
#pragma omp parallel
        {
            std::vector<Cell41> lab(18 * 18 * 18);
#pragma omp for schedule(dynamic, 4)
            for (int b = 0; b < NB; b++) {
                const int bx = b % nb, by = (b / nb) % nb, bz = b / (nb * nb);
                for (int z = -1; z <= 16; z++)
                    for (int y = -1; y <= 16; y++) {
                        for (int x = -1; x <= 16; x++) {
                            const int gx = (bx * 16 + x + N) % N, gy = (by * 16 + y + N) % N, gz = (bz * 16 + z + N) % N;
                            const int sb = (gx / 16) + nb * ((gy / 16) + nb * (gz / 16));
                            lab[(x + 1) + 18 * ((y + 1) + 18 * (z + 1))] =
                                grid[(size_t)sb * 4096 + (gx % 16) + 16 * ((gy % 16) + 16 * (gz % 16))];
                        }
                    }
                const int idx[3] = {bx, by, bz};
                const Real pos[3] = {0, 0, 0};
                const Real sp[3] = {(Real)h, (Real)h, (Real)h};
                BlockInfo info(b, idx, 0, pos, sp);
                LabD l = {lab.data()};
                BlockD o = {&grid[(size_t)b * 4096]};
                rhs(l, info, o);
            }
#pragma omp for schedule(static)
            for (int b = 0; b < NB; b++) {
                const int idx[3] = {0, 0, 0};
                BlockInfo info(b, idx, 0);
                BlockD o = {&grid[(size_t)b * 4096]};
                up(info, o);
            }
        }
    }
    for (int b = 0; b < NB; b++) {
        const int bx = b % nb, by = (b / nb) % nb, bz = b / (nb * nb);
        for (int i = 0; i < 4096; i++) {
            const size_t v = (size_t)(bx * 16 + i % 16) + N * ((size_t)(by * 16 + (i / 16) % 16) + (size_t)N * (bz * 16 + i / 256));
            phi[v] = grid[(size_t)b * 4096 + i].phi;
        }
    }
    return 0;
}
