// mrag_driver.cpp -- a forward driver written the way the reference's
// Glioma_HG_UQ is (Glioma/Glioma_HG_UQ.h:16-50, Glioma_HG_UQ.cpp:266-277 and
// :511-588): the member types, _reactionDiffusionStep and the time loop with
// adaptation keep the reference's calls, against include/gq_mrag.hpp instead
// of MRAG + TBB.  tests/test_mrag_mirror.py compiles it (CPU) and, on a GPU,
// runs it and compares its dump with the Python mirror's run of the same
// schedule.  Inputs: argv = bpd maxLevel order Dw rho tend dumpfreq tolr tolc
// out.bin; phi, pw, pg volumes (gpd^3 float32, x fastest) on stdin.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include "gq_mrag.hpp"

using namespace std;  // as the reference's MRAG headers do
using namespace gq::mrag;
typedef float Real;
#define _DIM 3
typedef Wavelets_Interp4thOrder W;
typedef Block<Real, 16, 16, 16> B;
struct RD_Projector_Wavelets {
};

class Glioma_HG_UQ {
private:
    Context ctx;
    Grid<W, B> *grid;
    BlockProcessing blockProcessing;
    BlockFWT<W, B, RD_Projector_Wavelets> blockfwt;
    Profiler profiler;
    int numberOfIterations = 0;
    double whenToWrite, whenToWriteOffset;
    bool bAdaptivity = true;
    int maxLevel;
    double refinement_tolerance, compression_tolerance;
    int blockSize = 16, blocksPerDimension;

    void _reactionDiffusionStep(BoundaryInfo *boundaryInfo, const int nParallelGranularity, const Real Dw,
                                const Real Dg, const Real rho, double dt)
    {
        vector<BlockInfo> vInfo = grid->getBlocksInfo();
        const BlockCollection<B> &collecton = grid->getBlockCollection();

        Glioma_ReactionDiffusionOperator<_DIM> rhs(Dw, Dg, rho);
        UpdateTumor<_DIM> updateTumor(dt);

        blockProcessing.pipeline_process(vInfo, collecton, *boundaryInfo, rhs);
        BlockProcessing::process(vInfo, collecton, updateTumor, nParallelGranularity);
    }

public:
    Glioma_HG_UQ(int bpd, int maxLevel_, double dumpfreq, double tolr, double tolc)
        : ctx(0), maxLevel(maxLevel_), refinement_tolerance(tolr), compression_tolerance(tolc), blocksPerDimension(bpd)
    {
        grid = new Grid<W, B>(ctx, bpd, maxLevel);
        whenToWriteOffset = dumpfreq;
        whenToWrite = whenToWriteOffset;
    }
    ~Glioma_HG_UQ() { delete grid; }
    Grid<W, B> &g() { return *grid; }

    void run(Real Dw, Real rho, double tend)
    {
        const int nParallelGranularity = (grid->getBlocksInfo().size() <= 8 ? 1 : 4);
        BoundaryInfo *boundaryInfo = &grid->getBoundaryInfo();
        Real Dg = 0.1 * Dw;
        double t = 0.0;
        double h = 1. / (blockSize * blocksPerDimension);
        double dt = 0.99 * h * h / (2. * _DIM * max(Dw, Dg));

        while (t <= tend) {
            _reactionDiffusionStep(boundaryInfo, nParallelGranularity, Dw, Dg, rho, dt);
            t += dt;
            numberOfIterations++;
            if (t >= ((double)(whenToWrite))) {
                if (bAdaptivity) {
                    Science::AutomaticRefinement<0, 0>(*grid, blockfwt, refinement_tolerance, maxLevel, 1, &profiler);
                    Science::AutomaticCompression<0, 0>(*grid, blockfwt, compression_tolerance, -1, &profiler);
                }
                whenToWrite = whenToWrite + whenToWriteOffset;
            }
        }
        // Refine final state & dump for UQ Likelihood
        if (bAdaptivity)
            Science::AutomaticRefinement<0, 0>(*grid, blockfwt, refinement_tolerance, maxLevel, 1, &profiler);
    }
    int iterations() const { return numberOfIterations; }
};

int main(int argc, char **argv)
{
    if (argc != 11) {
        std::fprintf(stderr, "usage: mrag_driver bpd maxLevel order Dw rho tend dumpfreq tolr tolc out.bin\n");
        return 2;
    }
    const int bpd = std::atoi(argv[1]), maxl = std::atoi(argv[2]);
    const Real Dw = (Real)std::atof(argv[4]), rho = (Real)std::atof(argv[5]);
    const double tend = std::atof(argv[6]), dumpfreq = std::atof(argv[7]);
    try {
        Glioma_HG_UQ sim(bpd, maxl, dumpfreq, std::atof(argv[8]), std::atof(argv[9]));
        const size_t gpd = (size_t)bpd * 16, n = gpd * gpd * gpd;
        std::vector<float> vol[3];
        for (int f = 0; f < 3; f++) {
            vol[f].resize(n);
            if (std::fread(vol[f].data(), sizeof(float), n, stdin) != n) throw std::runtime_error("short input");
            gq::mrag::check(sim.g().ctx(), gq_grid_write_volume(sim.g().ctx(), f, vol[f].data(), (int)gpd, 0),
                            "gq_grid_write_volume");
        }
        sim.run(Dw, rho, tend);
        std::vector<double> u = sim.g().dumpUniform();
        FILE *fo = std::fopen(argv[10], "wb");
        if (!fo || std::fwrite(u.data(), sizeof(double), u.size(), fo) != u.size()) throw std::runtime_error("write");
        std::fclose(fo);
        std::printf("steps %d blocks %d\n", sim.iterations(), sim.g().size());
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "mrag_driver: %s\n", e.what());
        return 1;
    }
}
