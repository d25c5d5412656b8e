// cli_likelihood.cpp -- gq_likelihood: drop-in for the reference's `likelihood`
// program (Likelihood/source/Likelihood_main.cpp:18-36, HGG_Likelihood.cpp:
// 11-189), evaluated on the GPU with FP64 accumulation.
//
// in : LikelihoodInput.txt (PETsigma2, PETscale, T1uc, T2uc, slope),
//      HGG_data.dat, PET_data.dat, T1_data.dat, T2_data.dat (Matrix D3D),
//      Sphere.dat (Matrix D2D, npoints x 3) when -selectivePoints 1
// out: Likelihood.txt with "-logL" printed as "%Lf \n"
// flags: -selectivePoints (1), -stepPet (3), -stepTi (2), -device (0)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/gq_mrag.hpp"
#include "matrix_io.hpp"

int main(int argc, const char **argv)
{
    auto flag = [&](const char *k, int d) {
        for (int i = 1; i + 1 < argc; i++)
            if (!std::strcmp(argv[i], k)) return std::atoi(argv[i + 1]);
        return d;
    };
    try {
        double prm[5];
        {
            std::ifstream f("LikelihoodInput.txt");
            if (!(f >> prm[0] >> prm[1] >> prm[2] >> prm[3] >> prm[4])) {
                std::printf("Aborting: missing input file LikelihoodInput.txt \n");
                return 1;
            }
        }
        const int selective = flag("-selectivePoints", 1);
        const int step_pet = flag("-stepPet", 3);
        const int step_ti = flag("-stepTi", 2);
        std::printf("PET: PETsigma2=%f, PETscale=%f \n", prm[0], prm[1]);
        std::printf("MRI: T1uc=%f, T2uc =%f, slope=%f \n", prm[2], prm[3], prm[4]);
        std::printf("stepPET=%d, stepTi =%d \n", step_pet, step_ti);
        gq::io::Volume model = gq::io::read_matrix("HGG_data.dat", 3);
        gq::io::Volume pet = gq::io::read_matrix("PET_data.dat", 3);
        gq::io::Volume t1 = gq::io::read_matrix("T1_data.dat", 3);
        gq::io::Volume t2 = gq::io::read_matrix("T2_data.dat", 3);
        const int n = model.n[0];
        if (model.n[1] != n || model.n[2] != n) throw std::runtime_error("HGG_data.dat must be a cube");
        for (const auto *v : {&pet, &t1, &t2})
            if (v->n[0] != n || v->n[1] != n || v->n[2] != n)
                throw std::runtime_error("observation size differs from HGG_data.dat");
        std::vector<int32_t> pts;
        if (selective) {
            gq::io::Volume sp = gq::io::read_matrix("Sphere.dat", 2);  // Points(i, c) = v[i + c*np]
            const int np = sp.n[0];
            if (sp.n[1] != 3) throw std::runtime_error("Sphere.dat must be npoints x 3");
            pts.resize((size_t)np * 3);
            for (int i = 0; i < np; i++)
                for (int c = 0; c < 3; c++) pts[(size_t)i * 3 + c] = (int32_t)sp(i, c);
        }
        gq::mrag::Context ctx(flag("-device", 0));
        gq::mrag::check(ctx.get(),
                        gq_obs_ref_upload(ctx.get(), n, pet.v.data(), t1.v.data(), t2.v.data(), pts.data(),
                                          (int)(pts.size() / 3)),
                        "gq_obs_ref_upload");
        double out[4];
        gq::mrag::check(ctx.get(),
                        gq_loglik_ref_volume(ctx.get(), model.v.data(), n, prm, selective, step_pet, step_ti, out),
                        "gq_loglik_ref_volume");
        std::printf("L_Pet=%f, L_T1=%f, L_T2 = %f \n", out[1], out[2], out[3]);
        std::printf("LogLike = %f \n", out[0]);
        FILE *f = std::fopen("Likelihood.txt", "w");
        if (!f) throw std::runtime_error("cannot write Likelihood.txt");
        std::fprintf(f, "%Lf \n", -(long double)out[0]);
        std::fclose(f);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "gq_likelihood: %s\n", e.what());
        return 1;
    }
}
