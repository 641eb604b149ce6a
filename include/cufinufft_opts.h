/* Options struct of the GPU NUFFT plan.
 *
 * Field order, types and meaning are those of the reference
 * (include/cufinufft_opts.h:7-40), so the struct is byte-compatible with the
 * Python ctypes mirror (python/cufinufft/cufinufft/_cufinufft.py:90-118) and
 * other bindings. gpu_stream holds a hipStream_t (an opaque pointer); 0 means
 * the null stream. Defaults are set by cufinufft_default_opts().
 *
 * gpu_method: 0 auto, 1 nonuniform-point driven (GM / GM-sort), 2 LDS
 * subproblem (SM), 3 output driven (OD; accepted and run by the SM kernels on
 * this GPU), 4 block-gather (not provided; rejected with
 * FINUFFT_ERR_METHOD_NOTVALID).
 */
#ifndef __CUFINUFFT_OPTS_H__
#define __CUFINUFFT_OPTS_H__

typedef struct cufinufft_opts {
  double upsampfac; /* upsampling ratio sigma; 0 = automatic */

  int gpu_method;
  int gpu_sort; /* method 1 only: 0 unsorted (GM), 1 sorted (GM-sort) */

  int gpu_binsizex; /* bin sizes of the sort / LDS tiles; 0 = automatic */
  int gpu_binsizey;
  int gpu_binsizez;

  int gpu_obinsizex; /* block-gather method only (unused here) */
  int gpu_obinsizey;
  int gpu_obinsizez;

  int gpu_maxsubprobsize;
  int gpu_kerevalmeth; /* 0: direct exp(sqrt()), 1: Horner polynomial */

  int gpu_spreadinterponly; /* 0: NUFFT, 1: spread or interpolate only */

  int gpu_maxbatchsize; /* vectors per FFT batch; 0 = automatic */

  int gpu_device_id;

  void *gpu_stream;

  int modeord; /* types 1,2: 0 CMCL order (-N/2..N/2-1), 1 FFT order */

  int gpu_np; /* output-driven points per pass; 0 = automatic */

  int debug; /* 0 silent, 1 timings and choices, 2 more */
} cufinufft_opts;

#endif
