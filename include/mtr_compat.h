/*
 * mtr_compat.h -- the reference's host interface, re-implemented over the
 * C ABI of mtr.h (C++, single process, no MPI).
 *
 * Same names, argument meaning and call order as Autumn1998/Multi-tilt-
 * Reconstuction, so its driver (MultiAxis.cpp) links against this header
 * with the MPI plumbing removed:
 *   struct Options            ReadAndInitialOptions.h:25-43 (+ sirt_iters, precision, form, device,
 *                             tol, checkpoint, resume)
 *   readOptions               ReadAndInitialOptions.cpp:127-172 (same short flags)
 *   struct InputData          MrcFileIO.h:13-29
 *   struct OutputModel        MrcFileIO.h:31-45
 *   struct DataToUse          MrcFileIO.h:47-53
 *   readMrcHeader             MrcFileIO.cpp:107-125
 *   readMrcAndTXBR            MrcFileIO.cpp:162-178
 *   updataAndWriteModel       MrcFileIO.cpp:194-201
 *   divideSlice               MultiAxis.cpp:13-31
 *   synchronizeVol            MultiAxis.cpp:73-99 (union box over agents, in process)
 *   ComputeInitialModel       FunctionOnGPU.cu:138-179 / FunctionOnGPU.cuh:41
 *   runSIRTOnGPU              FunctionOnGPU.cu:182-266 / FunctionOnGPU.cuh:42 (no MPI_Comm)
 *   initialZdata              FunctionOnGPU.cu:291-313 / FunctionOnGPU.cuh:43
 *   runMaceOnGPU              FunctionOnGPU.cu:336-362 / FunctionOnGPU.cuh:44 (Mann update,
 *                             arguments in their declared meaning)
 *   runMaceOnCPU              MultiAxis.cpp:33-38 (same arithmetic, runs on the GPU)
 *   reconstructModel          MultiAxis.cpp:40-71 (all agents in one process)
 * Error behaviour: where the reference prints and exits on a CUDA error,
 * these return a negative MTR_ERR_* code (success keeps the reference's
 * return value: 1 for ComputeInitialModel, 0 elsewhere) and leave the
 * message in mtr_last_error().
 */
#ifndef MTR_COMPAT_H_
#define MTR_COMPAT_H_

#include <string>
#include <vector>

#include "mtr.h"
#include "mtr_io.h"

namespace mtr_compat {

typedef mtr_volume_size VolumeSize;
typedef mtr_projection_size ProjectionSize;
typedef mtr_slice Slice;
typedef mtr_mrc_header_t MrcHeader;

constexpr int MAX_TILT = 16;

struct Options {
    std::vector<std::string> mrc_dir, txbr_dir;
    int number_tilt = 0;
    int block_x = 8, block_y = 8, block_z = 8;  // accepted; tiles are chosen by the library
    int iter_times = 5;                         // -n: MACE steps
    float sirt_step = 0.5f;                     // -s: lambda
    float mace_step = 0.7f;                     // -m: rho
    std::string out_dir;                        // -o
    int slice_num = 0;
    int sirt_iters = 10;                        // inner SIRT count (hard-coded 10 in the reference)
    int precision = MTR_PREC_FP32;
    int form = MTR_FORM_FAST;
    int device = 0;
    double tol = 0.0;           // --tol: stop once the consensus difference changes by <= tol (relative)
    std::string checkpoint;     // --checkpoint: write the MACE state here after the solve
    std::string resume;         // --resume: continue from this checkpoint instead of the BPT initial model
    int steps_done = 0;         // out: MACE steps done in total (resumed ones included)
    bool help = false;
};

struct InputData {
    std::vector<float> mrc_data;
    std::vector<double> x_coef, y_coef;
    MrcHeader mrc_header{};
    ProjectionSize prj{};
    VolumeSize vol{};
    int group_id = 0;
    int thread_id = 0;
    InputData();
};

struct OutputModel {
    MrcHeader output_header{};
    std::vector<float> output_data;
    VolumeSize vol{};
    Slice slc{};
    long long data_size = 0;
    int tilt_id = 0;
    void initialize(VolumeSize v);
};

struct DataToUse {
    int tilt_id = 0;
    int angN = 0;
    const float *mrc_data = nullptr;
    const double *x_coef = nullptr, *y_coef = nullptr;
};

int readOptions(int argc, char **argv, Options *opts);
void helpInfo();

int readMrcHeader(Options *opt, InputData *input_data);
int readMrcAndTXBR(Options *opt, InputData *input_data, DataToUse *local_data);
int updataAndWriteModel(Options *opts, OutputModel *model, const float *data);

Slice divideSlice(int myid, int N, int Z, int S);
void synchronizeVol(std::vector<InputData> &inputs);

int ComputeInitialModel(int device, DataToUse *local_data, InputData *input_data, OutputModel *model,
                        Options *opts);
int runSIRTOnGPU(int device, int leader_id, int group_id, DataToUse *local_data, OutputModel *model,
                 InputData *input_data, Options *opts);
int initialZdata(int device, OutputModel *model, Options *opts, float *z, const float *w,
                 const float *w_avg);
int runMaceOnGPU(int device, OutputModel *model, Options *opts, const float *v, float *w, const float *z,
                 float step);
int runMaceOnCPU(int myid, const float *v, float *w, const float *z, long long data_size, float step);
/* Whole reconstruction of every agent in one process: BPT-sum initial model
 * then opts->iter_times MACE steps; w_avg receives w-bar. */
int reconstructModel(std::vector<InputData> &inputs, Options *opts, std::vector<float> &w_avg,
                     double *seconds);

}  // namespace mtr_compat

#endif /* MTR_COMPAT_H_ */
