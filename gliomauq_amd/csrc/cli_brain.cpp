// cli_brain.cpp -- gq_brain: drop-in for the reference's `brain -anatomy HGG_UQ`
// forward solve (Glioma/Glioma_main.cpp:46-104, Glioma/Glioma_HG_UQ.cpp).
//
// File contract (same names and formats as the reference):
//   in : HGG_InputParameters.txt   Dw [cm^2/day], rho [1/day], T [days]   (:519-530)
//        HGG_TumorIC.bin           float32 x, y, z of the tumour centre     (:229-263)
//        <anatomyDir>/Patient%02d/P%02d_{GM,WM,CSF,T1,T2}.dat  (Matrix D3D) (:67-225)
//   out: HGG_data.dat              phi on the uniform (bpd*16)^3 grid       (:309-372)
// Flags: -adaptive -dumpfreq -verbose -pID as the reference, plus -anatomyDir,
// -bpd / -maxlevel / -wavelets (the reference's compile-time _BPD_,
// _MAXLEVEL_, _WAVELET_TYPE), -cfl (0.99), -tolref / -tolcomp (1e-4 / 1e-5,
// Glioma_Types.h:462-463), -device.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/gq_mrag.hpp"
#include "matrix_io.hpp"

namespace {

struct Args {
    int argc;
    const char **argv;
    const char *get(const char *k) const
    {
        for (int i = 1; i + 1 < argc; i++)
            if (!std::strcmp(argv[i], k)) return argv[i + 1];
        return nullptr;
    }
    double num(const char *k, double d) const
    {
        const char *v = get(k);
        return v ? std::atof(v) : d;
    }
    std::string str(const char *k, const char *d) const
    {
        const char *v = get(k);
        return v ? v : d;
    }
};

// Glioma_HG_UQ::_ic_PatientCase (Glioma_HG_UQ.cpp:67-225): anatomy volumes ->
// p_w, p_g per grid point (aspect-ratio-centred nearest-lower voxel, CSF
// removed under MRI signal, normalisation, CSF > 0.1 -> pure CSF with
// p_w = p_g = 0), and the smoothed-ball tumour of radius 0.005 with a
// transition over 2 points of h = 1/128.
void patient_anatomy(const gq::io::Volume &GM, const gq::io::Volume &WM, const gq::io::Volume &CSF,
                     const gq::io::Volume &T1, const gq::io::Volume &T2, const float tumor_ic[3],
                     const std::vector<gq_block_info> &info, std::vector<float> &phi, std::vector<float> &pw,
                     std::vector<float> &pg)
{
    const int sx = GM.n[0], sy = GM.n[1], sz = GM.n[2];
    const int smax = std::max(sx, std::max(sy, sz));
    const double bh = 1.0 / (double)smax;
    const float tumorRadius = 0.005f, smooth_sup = 2.f;
    const float h = 1.f / 128;
    const float iw = 1.f / (smooth_sup * h);
    const size_t nb = info.size();
    phi.assign(nb * 4096, 0.f);
    pw.assign(nb * 4096, 0.f);
    pg.assign(nb * 4096, 0.f);
    for (size_t b = 0; b < nb; b++) {
        const gq_block_info &bi = info[b];
        for (int iz = 0; iz < 16; iz++)
            for (int iy = 0; iy < 16; iy++)
                for (int ix = 0; ix < 16; ix++) {
                    // BlockInfo::pos: Real arithmetic, then widened (MRAGCommon.h:199-204)
                    const double x[3] = {(double)(bi.origin[0] + bi.h * ix), (double)(bi.origin[1] + bi.h * iy),
                                         (double)(bi.origin[2] + bi.h * iz)};
                    int m[3] = {(int)std::floor(x[0] / bh), (int)std::floor(x[1] / bh), (int)std::floor(x[2] / bh)};
                    m[0] -= (int)((smax - sx) * 0.5);
                    m[1] -= (int)((smax - sy) * 0.5);
                    m[2] -= (int)((smax - sz) * 0.5);
                    float PGt = 0, PWt = 0, Pcsf = 0, PT1 = 0, PT2 = 0;
                    if (!(m[0] < 0 || m[0] >= sx || m[1] < 0 || m[1] >= sy || m[2] < 0 || m[2] >= sz)) {
                        PGt = (float)GM(m[0], m[1], m[2]);
                        PWt = (float)WM(m[0], m[1], m[2]);
                        Pcsf = (float)CSF(m[0], m[1], m[2]);
                        PT1 = (float)T1(m[0], m[1], m[2]);
                        PT2 = (float)T2(m[0], m[1], m[2]);
                    }
                    const float MRIsignal = PT1 + PT2;
                    Pcsf = (MRIsignal > 0.) ? 0.f : Pcsf;
                    const double all = PWt + PGt + Pcsf;
                    const size_t c = b * 4096 + ix + 16 * (iy + 16 * iz);
                    if (all > 0) {
                        PGt = (float)(PGt / all);
                        PWt = (float)(PWt / all);
                        Pcsf = (float)(Pcsf / all);
                        Pcsf = ((double)Pcsf > 0.1) ? 1.f : Pcsf;
                        if (Pcsf < 1.f) {
                            pw[c] = PWt / (Pcsf + PWt + PGt);
                            pg[c] = PGt / (Pcsf + PWt + PGt);
                        }
                    }
                    const float p[3] = {(float)(x[0] - tumor_ic[0]), (float)(x[1] - tumor_ic[1]),
                                        (float)(x[2] - tumor_ic[2])};
                    const float dist = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
                    const float psi = (dist - tumorRadius) * iw;
                    if ((psi < -1) && (((double)PGt > 0.001) || ((double)PWt > 0.001))) phi[c] = 1.0f;
                    else if (((-1 <= psi) && (psi <= 1)) && ((PGt > 0) || (PWt > 0)))
                        phi[c] = (float)(1.0 * 0.5 * (1 - psi - std::sin(M_PI * psi) / (M_PI)));
                    else phi[c] = 0.f;
                }
    }
}

}  // namespace

int main(int argc, const char **argv)
{
    Args a{argc, argv};
    try {
        const bool verbose = a.num("-verbose", 0) != 0;
        const bool adaptive = a.num("-adaptive", 0) != 0;
        const int bpd = (int)a.num("-bpd", 8);
        const int maxlevel = (int)a.num("-maxlevel", bpd >= 16 ? 4 : 3);
        const int order = (int)a.num("-wavelets", 2);
        const double cfl = a.num("-cfl", 0.99);
        const double tolr = a.num("-tolref", 1e-4), tolc = a.num("-tolcomp", 1e-5);
        const double dumpfreq = a.num("-dumpfreq", 50);
        const int pID = (int)a.num("-pID", 0);
        const std::string dir = a.str("-anatomyDir", "../../Anatmoy/");

        // HGG_InputParameters.txt (read as Real = float, :519-530)
        float Dw_in = 0, rho_in = 0;
        double tend = 0;
        {
            std::ifstream f("HGG_InputParameters.txt");
            if (!(f >> Dw_in >> rho_in >> tend)) throw std::runtime_error("cannot read HGG_InputParameters.txt");
        }
        float tic[3] = {0, 0, 0};
        {
            FILE *f = std::fopen("HGG_TumorIC.bin", "rb");
            if (!f || std::fread(tic, 4, 3, f) != 3) throw std::runtime_error("cannot read HGG_TumorIC.bin");
            std::fclose(f);
        }
        char base[512];
        std::snprintf(base, sizeof(base), "%sPatient%02d/P%02d", dir.c_str(), pID, pID);
        auto rd = [&](const char *s) { return gq::io::read_matrix(std::string(base) + s, 3); };
        gq::io::Volume GM = rd("_GM.dat"), WM = rd("_WM.dat"), CSF = rd("_CSF.dat"), T1 = rd("_T1.dat"),
                       T2 = rd("_T2.dat");
        const int smax = std::max(GM.n[0], std::max(GM.n[1], GM.n[2]));
        const double L = smax * 0.1;  // voxel spacing 1 mm -> cm (:104)

        using namespace gq::mrag;
        Context ctx((int)a.num("-device", 0));
        Grid<> grid(ctx, bpd, maxlevel, order);
        std::vector<float> phi, pw, pg;
        patient_anatomy(GM, WM, CSF, T1, T2, tic, grid.getBlocksInfo(), phi, pw, pg);
        grid.write(GQ_FIELD_PHI, phi);
        grid.write(GQ_FIELD_PWM, pw);
        grid.write(GQ_FIELD_PGM, pg);

        // Glioma_HG_UQ::run (:511-588), in the reference's own calls
        BlockProcessing blockProcessing;
        BlockFWT<Wavelets_Interp4thOrder, Block16> blockfwt;
        Profiler profiler;
        BoundaryInfo *boundaryInfo = &grid.getBoundaryInfo();
        const int nParallelGranularity = grid.getBlocksInfo().size() <= 8 ? 1 : 4;
        float Dw = (float)(Dw_in / (L * L));
        float Dg = (float)(0.1 * Dw);
        const double h = 1. / (16 * bpd);
        const double dt = cfl * h * h / (2. * 3 * std::max(Dw, Dg));
        if (verbose) {
            std::printf("Dg=%e, Dw=%e, dt= %f, rho=%f , h=%f\n", Dg, Dw, dt, rho_in, h);
            std::printf("params Dw=%.9g Dg=%.9g rho=%.9g dt=%.17g\n", Dw, Dg, rho_in, dt);
        }
        double t = 0.0, when = dumpfreq;
        int iters = 0;
        while (t <= tend) {
            // _reactionDiffusionStep (:266-277)
            std::vector<BlockInfo> vInfo = grid.getBlocksInfo();
            const BlockCollection<Block16> &collection = grid.getBlockCollection();
            Glioma_ReactionDiffusionOperator<3> rhs(Dw, Dg, rho_in);
            UpdateTumor<3> updateTumor(dt);
            blockProcessing.pipeline_process(vInfo, collection, *boundaryInfo, rhs);
            BlockProcessing::process(vInfo, collection, updateTumor, nParallelGranularity);
            t += dt;
            iters++;
            if (t >= when) {
                if (adaptive) {
                    Science::AutomaticRefinement<0, 0>(grid, blockfwt, tolr, maxlevel, 1, &profiler);
                    Science::AutomaticCompression<0, 0>(grid, blockfwt, tolc, -1, &profiler);
                }
                when += dumpfreq;
                if (verbose) std::printf("Dumping data at time t=%f (blocks %d)\n", t, grid.size());
            }
        }
        if (adaptive) Science::AutomaticRefinement<0, 0>(grid, blockfwt, tolr, maxlevel, 1, &profiler);
        std::vector<double> u = grid.dumpUniform();
        const int gpd = bpd * 16;
        gq::io::write_matrix3d("HGG_data.dat", u.data(), gpd, gpd, gpd);
        if (verbose) std::printf("steps %d, blocks %d\n", iters, grid.size());
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "gq_brain: %s\n", e.what());
        return 1;
    }
}
