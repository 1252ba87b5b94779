// mtr_multiaxis: command-line driver with the reference's flags and file
// formats (MultiAxis.cpp:101-230 without MPI): reads K MRC tilt series and
// their TXBR geometry, reconstructs with BPT + MACE on one GPU and writes
// w-bar as a float MRC volume.
#include <cstdio>
#include <vector>

#include "mtr_compat.h"

using namespace mtr_compat;

int main(int argc, char **argv)
{
    Options opts;
    const int rc = readOptions(argc, argv, &opts);
    if (rc == 0) return 0;
    if (rc < 0 || opts.number_tilt <= 0 || opts.out_dir.empty()) {
        std::fprintf(stderr, "***WRONG INPUT.\n");
        helpInfo();
        return 2;
    }
    std::vector<InputData> inputs((size_t)opts.number_tilt);
    std::vector<DataToUse> local((size_t)opts.number_tilt);
    for (int a = 0; a < opts.number_tilt; a++) {
        inputs[a].group_id = a;
        inputs[a].thread_id = a;
        if (readMrcHeader(&opts, &inputs[a]) < 0) {
            std::fprintf(stderr, "***WRONG AT READ HEADER FILE (%s): %s\n", opts.mrc_dir[a].c_str(),
                         mtr_last_error());
            return 1;
        }
        if (readMrcAndTXBR(&opts, &inputs[a], &local[a]) < 0) {
            std::fprintf(stderr, "***Wrong INPUT (%s, %s)\n", opts.mrc_dir[a].c_str(), opts.txbr_dir[a].c_str());
            return 1;
        }
    }
    synchronizeVol(inputs);
    const VolumeSize &v = inputs[0].vol;
    std::printf("volume %d x %d x %d at (%d,%d,%d), %d tilt series, %d MACE steps x %d SIRT, lambda %g rho %g\n",
                v.X, v.Y, v.Z, v.Xstart, v.Ystart, v.Zstart, opts.number_tilt, opts.iter_times, opts.sirt_iters,
                opts.sirt_step, opts.mace_step);
    std::vector<float> w_avg;
    double secs = 0.0;
    const int sr = reconstructModel(inputs, &opts, w_avg, &secs);
    if (sr != MTR_OK) {
        std::fprintf(stderr, "reconstruction failed (%d): %s\n", sr, mtr_last_error());
        return 1;
    }
    std::printf("Total time: %.6f s (wall), MACE steps done: %d\n", secs, opts.steps_done);
    if (!opts.checkpoint.empty()) std::printf("Checkpoint: %s\n", opts.checkpoint.c_str());
    OutputModel res;
    res.slc = divideSlice(0, 1, v.Z, v.Zstart);
    res.initialize(v);
    if (updataAndWriteModel(&opts, &res, w_avg.data()) < 0) {
        std::fprintf(stderr, "***ERROR OCCURED WHEN WRITE MODEL.\n");
        return 1;
    }
    std::printf("Write finished: %s\n", opts.out_dir.c_str());
    return 0;
}
