/*
 * cv_oracle_avx.c -- AVX2 8-wide packet restatement of the reference path
 * (test infrastructure / CPU baseline only).  Filled in below.
 */
#include "cv_oracle.h"

int or_avx_available(void) { return 0; }

int or_avx_accumulate_frames(const or_scene* s, uint32_t width, uint32_t height, float* hdr,
                             uint32_t first_frame, uint32_t n_frames, uint32_t patch_begin,
                             uint32_t patch_end, int nthreads) {
    (void)s; (void)width; (void)height; (void)hdr; (void)first_frame; (void)n_frames;
    (void)patch_begin; (void)patch_end; (void)nthreads;
    return -1;
}
