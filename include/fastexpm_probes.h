/*
 * fastexpm measurement probes (libfastexpm_probes.so) -- not part of the
 * engine library and not on the expm path.
 *
 * No counterpart in the reference: these time single kernels of the engine
 * (and tilings it does not select) on device-resident pattern operands, the
 * sustained MFMA rates of the device, and a plain copy of known size whose
 * profiler counters calibrate FETCH_SIZE / WRITE_SIZE.  They back the
 * roofline and tiling figures that bench.py reports (--mfma-rate,
 * --copy-calibration).  Return codes are the FEXPM_* codes of fastexpm.h;
 * fexpm_probe_last_error() gives the calling thread's last failure message.
 */
#ifndef FASTEXPM_PROBES_H
#define FASTEXPM_PROBES_H

#include <stdint.h>

#include "fastexpm.h"

#ifdef __cplusplus
extern "C" {
#endif

const char* fexpm_probe_last_error(void);

/* Average device time (ms) of one GEMM launch over reps launches on
 * device-resident operands (m x k, n x k -> m x n).
 *   wtype 0 (FP32): cfg 0 the 128 x 128 default, 4 the 64 x 64 tiles,
 *     12 128 x 128 in 16 waves, 5..11 / 13..15 padding, wave-tile and
 *     MFMA-shape variants (mode 2 with sym);
 *   wtype 1 (FP64): cfg 0 4-wave 32x32, 1 8-wave 32x16, 2 128x64,
 *     3 64x128, 4 8-wave with 4-double row padding, 5 8-wave within 128
 *     VGPRs (2 and 3 are not square and are invalid with sym);
 *   wtype 2 / 3: the fp16x2 / bf16x3 split-precision squaring kernel alone
 *     (mode 2, sym); cfg % 10 the wave layout, cfg / 10 the tile group;
 *   wtype 4: the FP64 product by int8 slices, cfg = slices (4..6), 16 = six
 *     slices on 64 x 64 tiles. */
int fexpm_probe_gemm_time(int wtype, int mode, int sym, int cfg, int m, int n, int k, int reps, double* ms);

/* reps + 1 launches of a plain device copy of `bytes` (multiple of 16);
 * average ms of the last reps.  Reference traffic for the profiler's
 * FETCH_SIZE / WRITE_SIZE counters (each launch reads and writes bytes). */
int fexpm_probe_copy(int64_t bytes, int reps, double* ms);

/* Sustained MFMA rate of the device: a kernel whose waves issue only
 * independent MFMAs from registers (kind 0 FP32 16x16x4, 1 FP64 16x16x4,
 * 2 FP16 32x32x16, 3 BF16 32x32x16, 4 INT8 32x32x32 (TOP/s), 5 FP32
 * 32x32x2), 8 waves per SIMD; `iters` rounds per wave.  Returns TFLOP/s
 * and the kernel time: the practical peak that the GEMM kernels' rates are
 * compared with. */
int fexpm_probe_mfma_rate(int kind, int iters, double* tflops, double* ms);

/* The same probe with `waves` (1..8) waves per SIMD: the rate an MFMA stream
 * can reach at the occupancy of a register-heavy GEMM kernel (two waves per
 * SIMD for the squaring and int8-slice kernels). */
int fexpm_probe_mfma_rate_waves(int kind, int waves, int iters, double* tflops, double* ms);

#ifdef __cplusplus
}
#endif

#endif
