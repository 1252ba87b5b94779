// The reference's host interface (FunctionOnGPU.cuh, MrcFileIO.h,
// ReadAndInitialOptions.h, MultiAxis.cpp) over the C ABI, single process.
#include "mtr_compat.h"

#include <getopt.h>

#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>

namespace mtr_compat {

namespace {

constexpr int kEmptyLo = 999999, kEmptyHi = -999999;  // INF of MrcFileIOBaseFunction.h:20

struct Ctx {
    mtr_ctx *h = nullptr;
    ~Ctx()
    {
        if (h) mtr_ctx_destroy(h);
    }
};

VolumeSize slab_box(const OutputModel *m)
{
    VolumeSize v = m->vol;
    v.Zstart = m->slc.z_start;
    v.Zend = m->slc.z_end;
    v.Z = v.Zend - v.Zstart;
    return v;
}

int open_ctx(Ctx &c, int device, int precision, const VolumeSize &box)
{
    int rc = mtr_ctx_create(device, &c.h);
    if (rc != MTR_OK) return rc;
    if (precision != MTR_PREC_FP32 && (rc = mtr_ctx_set_precision(c.h, precision)) != MTR_OK) return rc;
    return mtr_ctx_set_volume(c.h, &box);
}

}  // namespace

InputData::InputData()
{
    vol.Xstart = vol.Ystart = vol.Zstart = kEmptyLo;
    vol.Xend = vol.Yend = vol.Zend = kEmptyHi;
    vol.X = vol.Y = vol.Z = 0;
}

void OutputModel::initialize(VolumeSize v)
{
    vol = v;
    data_size = (long long)vol.X * vol.Y * (slc.z_end - slc.z_start);
    output_data.assign((size_t)(data_size > 0 ? data_size : 0), 0.0f);
    std::memset(&output_header, 0, sizeof output_header);
    output_header.nx = vol.X;
    output_header.ny = vol.Y;
    output_header.nz = vol.Z;
    output_header.mode = MTR_MRC_FLOAT;
    output_header.nxstart = vol.Xstart;
    output_header.nystart = vol.Ystart;
    output_header.nzstart = vol.Zstart;
    output_header.mx = vol.X;
    output_header.my = vol.Y;
    output_header.mz = vol.Z;
}

void helpInfo()
{
    std::printf(
        "mtr_multiaxis -i K,st_1,...,st_K,txbr_1,...,txbr_K -o out.mrc [options]\n"
        "  -i K,...        K tilt series (MRC stacks) and their TXBR geometry files\n"
        "  -o FILE         output volume (MRC, float)\n"
        "  -n N            MACE steps (default 5)\n"
        "  -s LAMBDA       SIRT relaxation (default 0.5)\n"
        "  -m RHO          MACE Mann step (default 0.7)\n"
        "  -x/-y/-z N      accepted for compatibility; tiles are chosen by the library\n"
        "  --sirt-iters N  SIRT iterations per MACE step (default 10)\n"
        "  --precision P   fp32 | bf16 | fp16 storage (default fp32)\n"
        "  --form F        fast | plain (default fast)\n"
        "  --device D      GPU index (default 0)\n"
        "  --tune-cache F  keep the autotuner's tile choices in file F across runs\n"
        "  --tol T         stop early once the consensus difference changes by <= T (relative)\n"
        "  --checkpoint F  write the MACE state to F after the solve\n"
        "  --resume F      continue from checkpoint F (same inputs) instead of the initial model\n");
}

int readOptions(int argc, char **argv, Options *opts)
{
    static const struct option longopts[] = {
        {"help", no_argument, nullptr, 'h'},
        {"input", required_argument, nullptr, 'i'},
        {"output", required_argument, nullptr, 'o'},
        {"iters", required_argument, nullptr, 'n'},
        {"sirt-step", required_argument, nullptr, 's'},
        {"mace-step", required_argument, nullptr, 'm'},
        {"sirt-iters", required_argument, nullptr, 1001},
        {"precision", required_argument, nullptr, 1002},
        {"form", required_argument, nullptr, 1003},
        {"device", required_argument, nullptr, 1004},
        {"tune-cache", required_argument, nullptr, 1005},
        {"tol", required_argument, nullptr, 1006},
        {"checkpoint", required_argument, nullptr, 1007},
        {"resume", required_argument, nullptr, 1008},
        {nullptr, 0, nullptr, 0},
    };
    optind = 1;
    opterr = 0;
    int ch;
    while ((ch = getopt_long(argc, argv, ":hi:o:x:y:z:n:s:m:", longopts, nullptr)) != -1) {
        switch (ch) {
            case 'h': opts->help = true; helpInfo(); return 0;
            case 'x': opts->block_x = std::atoi(optarg); break;
            case 'y': opts->block_y = std::atoi(optarg); break;
            case 'z': opts->block_z = std::atoi(optarg); break;
            case 'n': opts->iter_times = std::atoi(optarg); break;
            case 's': opts->sirt_step = (float)std::atof(optarg); break;
            case 'm': opts->mace_step = (float)std::atof(optarg); break;
            case 'o': opts->out_dir = optarg; break;
            case 1001: opts->sirt_iters = std::atoi(optarg); break;
            case 1002:
                if (!std::strcmp(optarg, "fp32")) opts->precision = MTR_PREC_FP32;
                else if (!std::strcmp(optarg, "bf16")) opts->precision = MTR_PREC_BF16;
                else if (!std::strcmp(optarg, "fp16")) opts->precision = MTR_PREC_FP16;
                else { std::fprintf(stderr, "unknown precision '%s'\n", optarg); return -1; }
                break;
            case 1003:
                if (!std::strcmp(optarg, "fast")) opts->form = MTR_FORM_FAST;
                else if (!std::strcmp(optarg, "plain")) opts->form = MTR_FORM_PLAIN;
                else { std::fprintf(stderr, "unknown form '%s'\n", optarg); return -1; }
                break;
            case 1004: opts->device = std::atoi(optarg); break;
            // the library reads and extends the cache named by MTR_TUNE_CACHE
            case 1005: setenv("MTR_TUNE_CACHE", optarg, 1); break;
            case 1006: opts->tol = std::atof(optarg); break;
            case 1007: opts->checkpoint = optarg; break;
            case 1008: opts->resume = optarg; break;
            case 'i': {
                // K,mrc_1..mrc_K,txbr_1..txbr_K (ReadAndInitialOptions.cpp:104-125)
                std::stringstream ss(optarg);
                std::string tok;
                std::vector<std::string> parts;
                while (std::getline(ss, tok, ',')) parts.push_back(tok);
                if (parts.empty()) return -1;
                const int K = std::atoi(parts[0].c_str());
                if (K <= 0 || K > MAX_TILT || (int)parts.size() != 1 + 2 * K) {
                    std::fprintf(stderr, "-i expects K,mrc_1..mrc_K,txbr_1..txbr_K with 1 <= K <= %d\n", MAX_TILT);
                    return -1;
                }
                opts->number_tilt = K;
                opts->mrc_dir.assign(parts.begin() + 1, parts.begin() + 1 + K);
                opts->txbr_dir.assign(parts.begin() + 1 + K, parts.end());
                break;
            }
            case ':':
                std::fprintf(stderr, "Missing option argument for '%s'.\n", argv[optind - 1]);
                return -1;
            default:
                std::fprintf(stderr, "Invalid option '%s'.\n", argv[optind - 1]);
                return -1;
        }
    }
    if (opts->sirt_iters < 0 || opts->iter_times < 0) return -1;
    if (!(opts->tol >= 0.0)) {
        std::fprintf(stderr, "--tol must be >= 0\n");
        return -1;
    }
    if (opts->form == MTR_FORM_PLAIN && (opts->tol > 0.0 || !opts->checkpoint.empty() || !opts->resume.empty())) {
        std::fprintf(stderr, "--tol, --checkpoint and --resume run with --form fast only\n");
        return -1;
    }
    return 1;
}

int readMrcHeader(Options *opt, InputData *in)
{
    const int i = in->group_id;
    if (i < 0 || i >= (int)opt->mrc_dir.size()) return MTR_ERR_RANGE;
    int rc = mtr_mrc_read_header(opt->mrc_dir[i].c_str(), &in->mrc_header);
    if (rc != MTR_OK) {
        std::fprintf(stderr, "cannot read MRC header of %s\n", opt->mrc_dir[i].c_str());
        return rc;
    }
    in->prj.X = in->mrc_header.nx;
    in->prj.Y = in->mrc_header.ny;  // the reference uses nx here (MrcFileIO.cpp:121)
    in->prj.AngN = in->mrc_header.nz;
    return 1;
}

int readMrcAndTXBR(Options *opt, InputData *in, DataToUse *local)
{
    const int i = in->group_id;
    if (i < 0 || i >= (int)opt->txbr_dir.size()) return MTR_ERR_RANGE;
    const long long P = (long long)in->prj.X * in->prj.Y * in->prj.AngN;
    if (P <= 0) return MTR_ERR_STATE;
    in->mrc_data.resize((size_t)P);
    int rc = mtr_mrc_read_data(opt->mrc_dir[i].c_str(), in->mrc_data.data(), P);
    if (rc != MTR_OK) {
        std::fprintf(stderr, "cannot read MRC data of %s\n", opt->mrc_dir[i].c_str());
        return rc;
    }
    in->x_coef.assign(4 * (size_t)in->prj.AngN, 0.0);
    in->y_coef.assign(4 * (size_t)in->prj.AngN, 0.0);
    int views = 0;
    rc = mtr_txbr_read(opt->txbr_dir[i].c_str(), &in->vol, in->x_coef.data(), in->y_coef.data(), in->prj.AngN,
                       &views);
    if (rc != MTR_OK || views != in->prj.AngN) {
        std::fprintf(stderr, "cannot read TXBR %s (%d views for %d projections)\n", opt->txbr_dir[i].c_str(),
                     views, in->prj.AngN);
        return rc != MTR_OK ? rc : MTR_ERR_SHAPE;
    }
    local->tilt_id = i;
    local->angN = in->prj.AngN;
    local->mrc_data = in->mrc_data.data();
    local->x_coef = in->x_coef.data();
    local->y_coef = in->y_coef.data();
    return 1;
}

int updataAndWriteModel(Options *opts, OutputModel *model, const float *data)
{
    int rc = mtr_mrc_write_volume(opts->out_dir.c_str(), &model->vol, data);
    if (rc != MTR_OK) {
        std::fprintf(stderr, "cannot write %s\n", opts->out_dir.c_str());
        return rc;
    }
    return 1;
}

Slice divideSlice(int myid, int N, int Z, int S)
{
    Slice s{S, S};
    mtr_divide_slice(myid, N, Z, S, &s);
    return s;
}

void synchronizeVol(std::vector<InputData> &inputs)
{
    if (inputs.empty()) return;
    VolumeSize u = inputs[0].vol;
    for (auto &in : inputs) {
        u.Xstart = std::min(u.Xstart, in.vol.Xstart);
        u.Ystart = std::min(u.Ystart, in.vol.Ystart);
        u.Zstart = std::min(u.Zstart, in.vol.Zstart);
        u.Xend = std::max(u.Xend, in.vol.Xend);
        u.Yend = std::max(u.Yend, in.vol.Yend);
        u.Zend = std::max(u.Zend, in.vol.Zend);
    }
    u.X = u.Xend - u.Xstart;
    u.Y = u.Yend - u.Ystart;
    u.Z = u.Zend - u.Zstart;
    for (auto &in : inputs) in.vol = u;
}

int ComputeInitialModel(int device, DataToUse *local, InputData *in, OutputModel *model, Options *opts)
{
    (void)opts;
    Ctx c;
    int rc = open_ctx(c, device, MTR_PREC_FP32, slab_box(model));
    if (rc != MTR_OK) return rc;
    int id = 0;
    if ((rc = mtr_agent_add(c.h, &in->prj, local->x_coef, local->y_coef, local->mrc_data, &id)) != MTR_OK) return rc;
    if ((long long)model->output_data.size() < model->data_size) return MTR_ERR_SHAPE;
    if ((rc = mtr_bpt(c.h, id, model->output_data.data())) != MTR_OK) return rc;
    return 1;
}

int runSIRTOnGPU(int device, int leader_id, int group_id, DataToUse *local, OutputModel *model, InputData *in,
                 Options *opts)
{
    (void)leader_id;
    (void)group_id;
    // One process holds the whole box: the forward projection needs every
    // slab (the reference sums slab partials over MPI).
    if (model->slc.z_start != model->vol.Zstart || model->slc.z_end != model->vol.Zend) return MTR_ERR_SHAPE;
    Ctx c;
    int rc = open_ctx(c, device, opts->precision, slab_box(model));
    if (rc != MTR_OK) return rc;
    int id = 0;
    if ((rc = mtr_agent_add(c.h, &in->prj, local->x_coef, local->y_coef, local->mrc_data, &id)) != MTR_OK) return rc;
    rc = mtr_sirt(c.h, id, model->output_data.data(), opts->sirt_iters, opts->sirt_step, model->output_data.data());
    return rc;
}

int initialZdata(int device, OutputModel *model, Options *opts, float *z, const float *w, const float *w_avg)
{
    (void)opts;
    Ctx c;
    int rc = mtr_ctx_create(device, &c.h);
    if (rc != MTR_OK) return rc;
    const long long n = (long long)model->vol.X * model->vol.Y * model->vol.Z;
    return mtr_op_form_z(c.h, n, w, w_avg, z);
}

int runMaceOnGPU(int device, OutputModel *model, Options *opts, const float *v, float *w, const float *z,
                 float step)
{
    (void)opts;
    Ctx c;
    int rc = mtr_ctx_create(device, &c.h);
    if (rc != MTR_OK) return rc;
    const long long n = (long long)model->vol.X * model->vol.Y * model->vol.Z;
    return mtr_op_mann(c.h, n, v, w, z, step);
}

int runMaceOnCPU(int myid, const float *v, float *w, const float *z, long long data_size, float step)
{
    (void)myid;  // the reference's rank; the update runs on device 0
    Ctx c;
    int rc = mtr_ctx_create(0, &c.h);
    if (rc != MTR_OK) return rc;
    return mtr_op_mann(c.h, data_size, v, w, z, step);
}

int reconstructModel(std::vector<InputData> &inputs, Options *opts, std::vector<float> &w_avg, double *seconds)
{
    if (inputs.empty()) return MTR_ERR_STATE;
    Ctx c;
    int rc = open_ctx(c, opts->device, opts->precision, inputs[0].vol);
    if (rc != MTR_OK) return rc;
    for (auto &in : inputs) {
        int id = 0;
        rc = mtr_agent_add(c.h, &in.prj, in.x_coef.data(), in.y_coef.data(), in.mrc_data.data(), &id);
        if (rc != MTR_OK) return rc;
    }
    const VolumeSize &v = inputs[0].vol;
    w_avg.assign((size_t)v.X * v.Y * v.Z, 0.0f);
    const auto t0 = std::chrono::steady_clock::now();
    if (opts->tol <= 0.0 && opts->checkpoint.empty() && opts->resume.empty()) {
        rc = mtr_mace_solve(c.h, opts->iter_times, opts->sirt_iters, opts->sirt_step, opts->mace_step, opts->form,
                            w_avg.data());
        opts->steps_done = rc == MTR_OK ? opts->iter_times : 0;
    } else {
        // convergence control and checkpoints run the fast form only
        int before = 0, run = 0;
        rc = opts->resume.empty() ? mtr_mace_init(c.h) : mtr_checkpoint_load(c.h, opts->resume.c_str(), &before);
        if (rc == MTR_OK)
            rc = mtr_mace_continue(c.h, opts->iter_times, opts->sirt_iters, opts->sirt_step, opts->mace_step,
                                   opts->tol, &run, w_avg.data());
        opts->steps_done = before + run;
        if (rc == MTR_OK && !opts->checkpoint.empty())
            rc = mtr_checkpoint_save(c.h, opts->checkpoint.c_str(), opts->steps_done);
    }
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return rc;
}

}  // namespace mtr_compat
