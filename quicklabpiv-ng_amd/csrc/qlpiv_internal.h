// Shared declarations of the MI355X (gfx950) PIV correlation library.
//
// Every kernel of this library is compiled with -ffp-contract=off: the
// reference's Java arithmetic never fuses a multiply into an add, and the
// parity contract (integer peaks bit-exact, planes value-identical) depends on
// doing the same operations in the same order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/qlpiv.h"

namespace qlpiv {

// Twiddle tables, one per transform length M = 2N in {32, 64, 128, 256}, in the
// form (c[k], s[k]) consumed by the generic radix-2 butterfly
//   lo = ((a_r + c*b_r) - s*b_i, (a_i + c*b_i) + s*b_r)
//   hi = ((a_r - c*b_r) + s*b_i, (a_i - c*b_i) - s*b_r)
// for k in [0, M/2):
//   TW_FWD     : c = ws[k + M/4],         s = -ws[k]   (SimpleFFT.computeFFTStep :173-174)
//   TW_INV_CPU : c = ws[(M/4 - k) mod M], s =  ws[k]   (SimpleFFT.computeIFFTStep :300-307)
//   TW_INV_GPU : c = ws[k + M/4],         s =  ws[k]   (the kernel's re/im swap trick,
//                CrossCorrelationRealFFTParallelKernel.java:619,684, which is the
//                forward butterfly on swapped data)
// with ws[n] = (float)sin(2*pi*n/M) evaluated in double.
enum TwKind { TW_FWD = 0, TW_INV_CPU = 1, TW_INV_GPU = 2 };

constexpr int TW_SIZES = 4;
struct DeviceTables {
    float2 *tw[TW_SIZES][3];  // [log2M - 5][kind] -> M/2 entries
};

// Two implementations of the same correlation arithmetic: one lane per line of
// the transform (default) and the earlier LDS-exchange form (8/16 elements per
// lane, two LDS round trips per 1D transform).  QLPIV_XCORR_FORM=lds selects the
// latter when a context is created.
enum XcorrForm { FORM_LANE = 0, FORM_LDS = 1 };

struct Context {
    int device = 0;
    int xcorr_form = FORM_LANE;
    size_t hbm_budget = 0;
    size_t hbm_in_use = 0;
    hipStream_t stream = nullptr;
    DeviceTables tables{};
    // scratch buffers of the host-pointer entry points (grown on demand)
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // workspace of qlp_validate_vectors_dev (grown on demand)
    void *vwork = nullptr;
    size_t vwork_bytes = 0;
    // workspace of qlp_inherit_velocity_dev (grown on demand)
    void *iwork = nullptr;
    size_t iwork_bytes = 0;
    // per-work-group tile slots of the IA-128 correlation (grown on demand)
    void *bigwork = nullptr;
    size_t bigwork_bytes = 0;
    // MultiPeaks entry workspace (grown on demand; plans size it when created)
    void *mpwork = nullptr;
    size_t mpwork_bytes = 0;
    // constant DFT fragments of the MFMA correlation form, [N 16/32/64][fp16/bf16]
    void *mfma_tab[3][2] = {};
    // kernel shapes measured by the batch autotuner, by class key
    std::map<uint64_t, int> tuned_shape;
};

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
// Status of the exception being handled (call inside a catch block).
int fail_exception();
int check_hip(hipError_t e, const char *what);
Context *ctx_of(qlp_ctx *h);

// B-tile formation of a correlation pass whose tiles come from images.
enum BMode {
    B_CLIP = 0,        // window at (top, left)
    B_CLIP_SHIFT = 1,  // window at (round(top + u), round(left + v)): NoImageWarpingStrategy :79-81
    B_WARP_MICRO = 2,  // bilinear micro warp, SecondImage order
};

// Kernel shape of an exact correlation pass.  Every shape computes the same
// values; the batch path's autotuner picks among them.
enum PassShape {
    SHAPE_AUTO = 0,      // measured defaults (lane form for N = 16, else in-LDS)
    SHAPE_LANE = 1,      // one lane per transform line (N = 16 only)
    SHAPE_LDS = 2,       // one work-group per IA with the default work-group size
                         // (N = 128: 512 threads, 32-line batches)
    SHAPE_LDS_HALF = 3,  // half the work-group size, tiles input only
                         // (N = 128: 256 threads, 16-line batches)
    SHAPE_LDS_WIDE = 4,  // twice the work-group size, tiles input only (N <= 64)
};

// One correlation pass over n IA pairs.  Tiles come either from u8 tile arrays
// (tA/tB/maxA/maxB) or straight from two device-resident images (imgA/imgB with
// per-IA origins top/left and B displacement du/dv).  Outputs are optional:
// planes (n x D x D), FromCenter peaks (n x 3: i, j, value), min_floor (n) and
// the peak neighbourhood (n x 2 x (2*nbr_half+1): the column through the peak,
// then the row) used by the sub-pixel fits.
struct PassParams {
    int n = 0, N = 0, mode = QLP_NORM_PER_TILE;
    const uint8_t *tA = nullptr, *tB = nullptr;
    const float *tAf = nullptr, *tBf = nullptr;  // float tiles (MatrixFloat), in place of tA / tB
    const float *maxA = nullptr, *maxB = nullptr;
    const uint8_t *imgA = nullptr, *imgB = nullptr;
    int H = 0, W = 0;
    int per_pair = 0;  // IAs per image pair (0: all n IAs belong to one pair);
                       // pair k uses images imgA/imgB + k*H*W and origins top/left[ia % per_pair]
    const int *top = nullptr, *left = nullptr;
    const float *du = nullptr, *dv = nullptr;
    int bmode = B_CLIP;
    bool bmax255 = false;  // B tiles carry the maximum 255 (cut from a warped image:
                           // a new MatrixByte region, MatrixByte.java:40-45, 366-384)
    float *planes = nullptr;
    float *peaks = nullptr, *min_floor = nullptr;
    float *nbr = nullptr;
    int nbr_half = 0;
    int arith = QLP_XCORR_EXACT;  // QLP_XCORR_*: arithmetic form of the correlation
    int shape = 0;                // PassShape of an exact pass
    // MultiPeaks entries (an exact pass from the LDS plane, no fused peak):
    // per IA mp_geom(D).stride entries at mp_ent, chunk maxima at mp_cmax, the
    // plane minimum at mp_min (pass_writes_entries tells which passes do)
    uint2 *mp_ent = nullptr;
    float *mp_cmax = nullptr, *mp_min = nullptr;
    // Persistent launch: the IA count lives in device memory (written before the
    // launch, e.g. by a step of a captured graph) and n is only the capacity of
    // the buffers.  The grid is then sized to the work-groups the device holds
    // at once and each work-group loops over IAs.
    const int *n_dev = nullptr;
};

// Work-groups of one kernel that the device holds at once.
template <typename K>
static inline int resident_groups(K kernel, int threads, size_t lds = 0)
{
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 256;
    return (per_cu < 1 ? 1 : per_cu) * cus;
}

// Grid of a pass over n IAs in blocks of 'per_block' IAs: one block per group
// of IAs, or for a persistent launch at most the resident work-groups.
static inline int pass_grid(const PassParams &p, int per_block, int resident)
{
    const int blocks = (p.n + per_block - 1) / per_block;
    if (!p.n_dev) return blocks;
    return blocks < resident ? (blocks < 1 ? 1 : blocks) : resident;
}

// Launchers (qlpiv_xcorr.hip, qlpiv_peaks.hip, qlpiv_subpixel.hip, qlpiv_warp.hip).
int launch_xcorr(Context *ctx, int n, int N, const uint8_t *tA, const uint8_t *tB,
                 const float *maxA, const float *maxB, int mode, float *planes);
int launch_pass(Context *ctx, const PassParams &p);
// Whether launch_pass(ctx, p) runs a kernel that writes p.mp_* (the exact pass
// with its plane in LDS).
bool pass_writes_entries(const Context *ctx, const PassParams &p);
// The allocation-free correlation: the IA-128 slots in a static device array
// whose per-device address init_noalloc_scratch (with the device current)
// looks up once; launch_xcorr_noalloc allocates nothing.
int init_noalloc_scratch(int device);
int launch_xcorr_noalloc(Context *ctx, int n, int N, const uint8_t *tA, const uint8_t *tB,
                         const float *maxA, const float *maxB, int mode, float *planes);
int launch_pass_lane(Context *ctx, const PassParams &p, bool from_image, bool peak);
// IA 128: the tile lives in an HBM workspace (ctx->bigwork) of one 512 KB slot
// per resident work-group.  big_work_need is the workspace of a pass of n IAs
// with the given kernel (resident work-groups from the occupancy of this
// device); prepare_big grows the context's workspace to it; big_work_estimate
// is what n IAs need on a 256-CU device with one resident work-group per CU.
size_t big_work_need(int n, int mode, bool from_image, bool peak, int shape);
int prepare_big(Context *ctx, int n, int mode, bool from_image, bool peak, int shape);
size_t big_work_estimate(int n);
int ensure_bigwork(Context *ctx, size_t bytes);
// MultiPeaks entries form (kernel width 3, at most 4 peaks): each plane's block
// maxima in visit order in a workspace (entries, then chunk maxima), replayed
// with one lane per plane.  multipeak_work_bytes: the workspace of n planes of
// D x D (0 where the form does not apply); multipeak_work_carve: its two parts;
// launch_multipeak_replay: the replay (peaks, count) from the entries.
bool multipeak_entries_form(int D, int n_peaks, int kernel_width);
size_t multipeak_work_bytes(int n, int D, int n_peaks, int kernel_width);
void multipeak_work_carve(void *work, int n, int D, uint2 *&ent, float *&cmax);
int launch_multipeak_replay(Context *ctx, int n, int D, int P, const uint2 *ent,
                            const float *cmax, float *peaks, int *count);
int ensure_mpwork(Context *ctx, size_t bytes);
// Whether the workspace holds 'bytes' or can grow to it within the HBM budget
// (the MultiPeaks callers take the segment form, which needs none, otherwise).
bool mpwork_fits(const Context *ctx, size_t bytes);
// The throughput form (qlpiv_xcorr_mfma.hip): form is QLP_XCORR_MFMA_*.
int launch_pass_mfma(Context *ctx, const PassParams &p, int form, bool from_image, bool peak);
int launch_find_peaks(Context *ctx, int kind, int n, int D, const float *planes, int n_peaks,
                      int kernel_width, float *peaks, float *min_floor, int *count = nullptr);
// u, v (Guo only, or null): the fit kernel also applies the multi-pass loop's
// displacement update from each IA's main peak (as launch_accumulate).
int launch_subpixel(Context *ctx, int kind, int pixels, int iterations, int n, int D, int n_peaks,
                    const float *planes, const float *min_floor, const float *peaks_in,
                    float *peaks_out, float *u = nullptr, float *v = nullptr);
// n_dev: persistent launch (n is the capacity, the count is in device memory);
// fit_on: per-IA flag, 0 keeps the integer peak (status SKIPPED).
struct FitScatter;
// scatter (optional, with n_dev): each IA's updated displacement and fit record
// go straight to its pair level's arrays (the step engine's store)
int launch_subpixel_nbr(Context *ctx, int kind, int pixels, int iterations, int n, int D,
                        int nbr_half, const float *nbr, const float *min_floor,
                        const float *peaks_in, float *peaks_out, float *du, float *dv,
                        const int *n_dev = nullptr, const uint8_t *fit_on = nullptr,
                        const FitScatter *scatter = nullptr);
int launch_clip(Context *ctx, const uint8_t *img, int H, int W, int n, int N, const int *top,
                const int *left, uint8_t *tiles, float *maxv);
// Tile warps of one side (WarpSide, qlpiv_device.h; SIDE_B_SECOND is the
// SecondImage B tile).
int launch_warp_micro(Context *ctx, const uint8_t *img, int H, int W, int n, int N, const int *top,
                      const int *left, const float *u, const float *v, uint8_t *tiles,
                      float *maxv, int side = 0);
// Tiles of float frames (ImageFloat): a clip (at (top, left), or shifted by the
// rounded (u, v) when u and v are given) with the MatrixFloat maximum, and the
// Micro warp of one side with the maximum 65535.
// clamp16: the MatrixInt16 maximum (>= 16) instead of the MatrixFloat one (0 -> 1);
// wmax: the maximum of a warped tile (65535 for ImageFloat, 2^bits - 1 for ImageInt16).
int launch_clip_f32(Context *ctx, const float *img, int H, int W, int n, int N, const int *top,
                    const int *left, const float *du, const float *dv, float *tiles, float *maxv,
                    int clamp16 = 0);
int launch_warp_micro_f32(Context *ctx, const float *img, int H, int W, int n, int N, const int *top,
                          const int *left, const float *u, const float *v, float *tiles, float *maxv,
                          int side = 0, float wmax = 65535.0f);
int launch_u16_to_f32(Context *ctx, const uint16_t *in, float *out, size_t n);
// The Mini warp of one side of float frames (pairs x H x W) into float tiles
// with the maximum wmax (qlpiv_splinewarp.hip).
int launch_warp_mini_f32(Context *ctx, int pairs, const float *imgs, int H, int W, int nI, int nJ,
                         int N, const int *top, const int *left, const float *u, const float *v,
                         float *tiles, float *maxv, int side, float wmax);
// The Full warp of float frames (as launch_full_warp), and the maximum of each
// pair's frame (0 -> 1) into per_pair and spread to the pair's n tiles in dst.
int launch_full_warp_f32(Context *ctx, int pairs, const float *const *imgs, int H, int W, int nI,
                         int nJ, int tH, int tW, const int *top_h, const int *left_h, const float *u,
                         const float *v, void *work, float *const *outs, const int *sides,
                         int n_sides);
int launch_frame_max_spread(Context *ctx, const float *imgs, int pairs, int H, int W, int n,
                            float *per_pair, float *dst);
// Spline-field warps (qlpiv_splinewarp.hip).
int launch_warp_mini(Context *ctx, int pairs, const uint8_t *imgs, int H, int W, int nI, int nJ,
                     int N, const int *top, const int *left, const float *u, const float *v,
                     uint8_t *tiles, float *maxv, int side = 0);
// warpAndClipImage of the Micro / Mini strategies at a level > 0
// (BiLinearImageMicroAndMiniWarpingStrategy.java:39-186): the A and B tiles
// and maxima of the nI x nJ IAs of 'pairs' frame pairs in warp order 'order'
// (QLP_WARP_ORDER_*): a warped tile carries the maximum 255, a clipped one its
// own (>= 16).  mode: QLP_WARP_MICRO or QLP_WARP_MINI.
int warp_clip_tiles(Context *ctx, int mode, int order, const uint8_t *imgsA, const uint8_t *imgsB,
                    int H, int W, int pairs, int nI, int nJ, int N, const int *top, const int *left,
                    const float *u, const float *v, uint8_t *tA, float *mA, uint8_t *tB, float *mB);
size_t full_warp_work_bytes(int pairs, int nI, int nJ);
// The field is built once; 'sides' lists the frames to warp with it
// (outs[k] receives imgs[k] warped for side sides[k]).
int launch_full_warp(Context *ctx, int pairs, const uint8_t *const *imgs, int H, int W, int nI,
                     int nJ, int tH, int tW, const int *top_h, const int *left_h, const float *u,
                     const float *v, void *work, uint8_t *const *outs, const int *sides,
                     int n_sides);

// Vector validation (qlpiv_validate.hip).  The device arrays of one launch:
// inputs, outputs and the workspace carved by valid_work_carve.
struct ValidLaunch {
    int pairs = 0, nI = 0, nJ = 0, P = 1;
    const int *top = nullptr, *left = nullptr, *rank = nullptr;  // nI*nJ
    float *u = nullptr, *v = nullptr;                             // pairs*n, in/out
    const float *pk = nullptr;  // peak displacement increments: (u_k, v_k) =
    int pk_stride = 2;          //   pk[(t*P + k)*pk_stride + {0,1}] - pk_off
    float pk_off = 0.0f;
    const int *pcount = nullptr;  // pairs*n or null
    uint8_t *invalid = nullptr;   // pairs*n or null
    int *stats = nullptr;         // pairs*4 or null
    float *bu = nullptr, *bv = nullptr;
    int *flags = nullptr, *done = nullptr, *act = nullptr, *act_count = nullptr;
    // act_count is 0 on entry (zeroed when its workspace was allocated; every
    // sweep returns it to 0): the launch then skips clearing it
    bool counts_zero = false;
};
int check_valid_config(const qlp_valid_config &cfg);
void spiral_rank(int nI, int nJ, int *rank);
size_t valid_work_bytes(int pairs, int n);
void valid_work_carve(void *base, int pairs, int n, ValidLaunch &vl);
int launch_validate(Context *ctx, const qlp_valid_config &cfg, const ValidLaunch &vl);
int launch_accumulate(Context *ctx, int n, int P, int D, const float *sub, float *u, float *v);

// Velocity inheritance (qlpiv_inherit.hip, qlpiv_pipeline.hip).  A tile grid:
// nI x nJ tiles of H x W pixels at host-side origins top/left (row-major).
struct InheritGrid {
    int nI, nJ, H, W;
    const int *top, *left;
};
int build_inherit_stencil(int mode, const InheritGrid &pg, const InheritGrid &cg,
                          std::vector<int> &idx, std::vector<float> &w, int &K);
int build_inherit_bicubic(const InheritGrid &pg, const InheritGrid &cg, std::vector<double> &ys,
                          std::vector<double> &xs, std::vector<float> &cy, std::vector<float> &cx);
int launch_inherit_stencil(Context *ctx, int pairs, int cn, int pn, int K, const int *idx,
                           const float *w, int round_it, const float *pu, const float *pv, float *u,
                           float *v);
int launch_inherit_bicubic(Context *ctx, int pairs, int pnI, int pnJ, int cn, const double *ys,
                           const double *xs, const float *cy, const float *cx, int round_it,
                           const float *pu, const float *pv, float *u, float *v);
int launch_inherit_direct(Context *ctx, int pairs, int nI, int nJ, int pI, int pJ, int round_it,
                          const float *pu, const float *pv, float *u, float *v);
int ensure_iwork(Context *ctx, size_t bytes);

// ---- the step engine (qlpiv_engine.hip): one adaptive level of many frame
// pairs per step, in one persistent, graph-captured launch -----------------
constexpr int ENG_CLASSES = 4;     // IA sizes 16, 32, 64, 128
constexpr int ENG_MAX_SEGS = 4096; // pair levels in one step, all classes
constexpr int ENG_STAGES = 4;      // steps whose tables may be queued at once
enum { STEP_CLIP = 0, STEP_SHIFT = 1, STEP_MICRO = 2 };
// One pair level of a step.
struct StepSeg {
    const uint8_t *a, *b;   // the pair's frames (device, H x W)
    const int *top, *left;  // the level's IA origins (device, nI * nJ)
    const float *pu, *pv;   // the parent level's field (device, pnI * pnJ), null at the first level
    float *u, *v;           // this level's field (device, nI * nJ), written by the step
    float *sub;             // fit records (device, nI * nJ x 4) or null
    int H, W, nI, nJ, pnI, pnJ;
    int first;              // first IA of the segment in its class (set by engine_step)
    int warp;               // B tile: STEP_CLIP, STEP_SHIFT (no warp, rounded inheritance)
                            // or STEP_MICRO (micro warp); the first level always clips
    int fit;                // 1: sub-pixel fit; 0: the integer peak (before the start step)
    int pad;
};
// Where a step's fits land: IA ia belongs to segment segs[seg_of[ia]] at
// index ia - first of its level arrays (u, v and, when set, sub).
struct FitScatter {
    const StepSeg *segs = nullptr;
    const int *seg_of = nullptr;
};
struct StepHeader {
    int n[ENG_CLASSES];     // IAs of each class in the step
    int nseg[ENG_CLASSES];  // segments of each class
    int base[ENG_CLASSES];  // first segment of each class
    int pad;
};
struct EngineConfig {
    int norm_mode, subpix_kind, subpix_pixels, subpix_iterations, xcorr_form;
};
struct Engine;
int eng_class_of(int ia);
size_t engine_bytes(const EngineConfig &cfg, const size_t cap[ENG_CLASSES]);
// cap: the most IAs of each class one step may hold (0: the class is unused).
int engine_create(Context *ctx, const EngineConfig &cfg, const size_t cap[ENG_CLASSES], Engine **out);
int engine_fix(Engine *e, int ia, int shape, int schedule);
// Enqueue one step on the context's stream (the first step tunes and builds
// the graph, synchronising).  segs[c]: the pair levels of class c.
int engine_step(Engine *e, const std::vector<StepSeg> segs[ENG_CLASSES]);
int engine_report(const Engine *e, qlp_batch_report *r);
size_t engine_cap(const Engine *e, int cls);
void engine_destroy(Engine *e);

// First-level IA origins of a frame (AdaptiveInterAreaStrategySuperPosition,
// qlpiv_pipeline.hip); margins = top, bottom, left, right.
void grid_origins(int H, int W, int tile, float overlap_factor, const int margins[4], int &nI,
                  int &nJ, std::vector<int> &top, std::vector<int> &left);
// IA origins of any adaptive level (first: the first level's placement, else
// the later levels' float accumulation), as a plan lays them out.
void level_origins(int H, int W, int tile, float overlap_factor, const int margins[4], bool first,
                   int &nI, int &nJ, std::vector<int> &top, std::vector<int> &left);
__global__ void fill_kernel(float *__restrict__ x, float value, size_t n);

}  // namespace qlpiv
