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
