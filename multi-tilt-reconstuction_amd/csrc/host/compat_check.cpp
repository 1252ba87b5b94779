// mtr_compat_check: drives each reference-shaped host function once on the
// files named on the command line and writes every result as an MRC volume,
// so a test can compare them with the CPU oracle.
//   mtr_compat_check st txbr outdir lambda sirt_iters rho
// writes outdir/{bpt,bpt_slab0,bpt_slab1,sirt,z,mann}.mrc
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "mtr_compat.h"

using namespace mtr_compat;

static int write(const std::string &path, const VolumeSize &v, const float *d)
{
    return mtr_mrc_write_volume(path.c_str(), &v, d) == MTR_OK ? 0 : 1;
}

int main(int argc, char **argv)
{
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s st txbr outdir lambda sirt_iters rho\n", argv[0]);
        return 2;
    }
    Options opts;
    opts.mrc_dir = {argv[1]};
    opts.txbr_dir = {argv[2]};
    opts.number_tilt = 1;
    opts.sirt_step = (float)std::atof(argv[4]);
    opts.sirt_iters = std::atoi(argv[5]);
    const float rho = (float)std::atof(argv[6]);
    const std::string out = argv[3];
    InputData in;
    DataToUse local;
    if (readMrcHeader(&opts, &in) < 0 || readMrcAndTXBR(&opts, &in, &local) < 0) return 1;
    const VolumeSize v = in.vol;
    // BPT of the whole box, then of two slabs (divideSlice) stitched.
    OutputModel m;
    m.slc = divideSlice(0, 1, v.Z, v.Zstart);
    m.initialize(v);
    if (ComputeInitialModel(0, &local, &in, &m, &opts) != 1) return 1;
    std::vector<float> bpt = m.output_data;
    if (write(out + "/bpt.mrc", v, bpt.data())) return 1;
    std::vector<float> stitched;
    for (int r = 0; r < 2; r++) {
        OutputModel s;
        s.slc = divideSlice(r, 2, v.Z, v.Zstart);
        s.initialize(v);
        if (ComputeInitialModel(0, &local, &in, &s, &opts) != 1) return 1;
        stitched.insert(stitched.end(), s.output_data.begin(), s.output_data.end());
    }
    if (write(out + "/bpt_slabs.mrc", v, stitched.data())) return 1;
    // SIRT from the BPT
    if (runSIRTOnGPU(0, 0, 0, &local, &m, &in, &opts) != MTR_OK) return 1;
    std::vector<float> sirt = m.output_data;
    if (write(out + "/sirt.mrc", v, sirt.data())) return 1;
    // z = 2 wbar - w with w = BPT, wbar = SIRT result
    std::vector<float> z(bpt.size());
    if (initialZdata(0, &m, &opts, z.data(), bpt.data(), sirt.data()) != MTR_OK) return 1;
    if (write(out + "/z.mrc", v, z.data())) return 1;
    // Mann: w = rho (2 v - z) + (1 - rho) w with v = SIRT result, w = BPT
    std::vector<float> w = bpt;
    if (runMaceOnGPU(0, &m, &opts, sirt.data(), w.data(), z.data(), rho) != MTR_OK) return 1;
    if (write(out + "/mann.mrc", v, w.data())) return 1;
    std::vector<float> w2 = bpt;
    if (runMaceOnCPU(0, sirt.data(), w2.data(), z.data(), (long long)w2.size(), rho) != MTR_OK) return 1;
    if (write(out + "/mann_cpu_name.mrc", v, w2.data())) return 1;
    std::printf("ok\n");
    return 0;
}
