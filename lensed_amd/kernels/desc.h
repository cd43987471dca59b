// Per-system launch descriptor shared by the host runtime and the device
// kernels.  One descriptor per observed system in a launch; the kernels walk
// a flat list of work items and find their system by the item ranges.
#pragma once

#define LENSED_RENDER_BS 256  // threads (= pixels) per render work item: a 16 x 16 block,
#define LENSED_RENDER_SIDE 16 // an 8 x 8 square of it per wave
#define LENSED_CULL_GROUP 64  // render items per pixel-cull work item: 64 squares to each of its 4 waves
#define LENSED_LIST_ENTRY 16  // 4 x 4 quads per entry of the render list: a render block, four to a wave
#define LENSED_CONV_TX 64     // conv tile width  (threads along x)
#define LENSED_CONV_TY 4      // threads along y
#define LENSED_CONV_RY 8      // output rows per thread, large launches (a multiple of 8)
#define LENSED_CONV_RY_SMALL 2  // output rows per thread, launches too small to fill the GPU
#define LENSED_CONV_RY_TALL 16  // output rows per thread of the 64-row tile (a tuning choice)
#define LENSED_CONV_SWEEP 1   // conv_rows value of the column sweep: one wave per 64-column strip
#define LENSED_SWEEP_MAX_H 512  // the sweep's 8-row band flags fit one 64-bit mask
#define LENSED_SWEEP_MAX_PSF 21  // the sweep's PSF sides at most: its PH-row window is unrolled, and its
                                // 2 x 21 taps stay in scalar registers (a 31-row window is no longer
                                // unrolled by the compiler and indexes registers indirectly)
// chi^2 partial sums are formed over fixed 8-row x 64-column blocks, so the
// short tile (4 x 2 rows) is exactly one block, the default one (4 x 8) four
// and the tall one (4 x 16) eight
#define LENSED_CONV_BS (LENSED_CONV_TX * LENSED_CONV_TY)
#define LENSED_PSF_MAX 63     // largest PSF side the conv tile is sized for
#define LENSED_EPL_NT 256     // EPL angular table: intervals of the diamond angle p in [0, 1]
#define LENSED_EPL_ROWS (LENSED_EPL_NT + 1)  // rows per component: the intervals and the end point p = 1
#define LENSED_EPL_TABLE_F4 (2 * LENSED_EPL_ROWS)  // float4 per candidate (2 components)
// float4 index of row i of component comp (0: x, 1: y): each component's
// rows are contiguous (16-byte stride), so lanes reading nearby rows hit
// different LDS banks
#define LENSED_EPL_AT(comp, i) ((comp) * LENSED_EPL_ROWS + (i))
// per candidate in npre[]: the nq node pairs, then the pixel cull's constants
#define LENSED_NPRE_CULL 3
#define LENSED_NPRE_STRIDE(nq) ((nq) + LENSED_NPRE_CULL)

struct lensed_desc {
    int width, height, npix, model;  // image size, pixels, model index
    int psfw, psfh, ncand, npar;     // PSF size (0 = none), candidates, params per candidate
    int ntile_r, ntile_c, tiles_x, dwords;  // render tiles, conv tiles, conv tiles along x, block words
    float pcs[4];                    // pixel coordinates (rx, ry, sx, sy)
    long long image_off;             // floats into image[] / weight[]
    long long psf_off;               // floats into psf[]
    long long model_off;             // floats into model[]  ([ncand][npix])
    long long block_off;             // words into blocks[]  ([ncand][dwords])
    long long param_off;             // floats into params[] ([ncand][npar])
    long long partial_off;           // doubles into partials[] ([ncand][ntile_c])
    long long out_off;               // doubles into logl[]  ([ncand])
    long long table_off;             // float2 into the EPL angular tables ([ncand][...])
    int item_r, item_c, item_s, item_f;  // first work item in render, conv, set, final
    int item_t, epl_obj, epl_q, epl_t;   // table items; EPL object word offset, q and t param index
    int quad_off, nq, conv_ry, sep_rank;  // nodes (float4 offset, count); conv rows/thread; PSF rank (0 = direct)
    long long sep_off;                   // floats into psf[]: per rank, column [psfh] then row [psfw] factor
    int npart, table_mode, item_e, item_k;  // chi^2 partials per candidate (8-row x 64-column blocks); EPL table
                                            // in use; first emulator node item (NT + 1 nodes per candidate);
                                            // first cull item (LENSED_CULL_GROUP render items each)
    long long psft_off;                  // floats into psf[]: the taps transposed, [psfw][psfh]
    long long npre_off;                  // float4 into npre[]: node pairs, cull constants ([ncand][nq + 3])
    float qrad;                          // largest distance of a quadrature node from the pixel centre
    int cull;                            // skip pixels whose rays provably all miss the source
    long long dark_off;                  // doubles into the data's chi^2 block partials for a model of 0 ([npart])
};

// LDS row stride of the conv tile: 64 + PW - 1 columns plus room for the
// vector reads of the row pass, a multiple of 4 floats
#define LENSED_TILE_STRIDE(PW) (((LENSED_CONV_TX + (PW) - 1 + 3) + 3) / 4 * 4)
