// Per-system launch descriptor shared by the host runtime and the device
// kernels.  One descriptor per observed system in a launch; the kernels walk
// a flat list of work items and find their system by the item ranges.
#pragma once

#define LENSED_RENDER_BS 256  // threads (= pixels) per render work item
#define LENSED_CONV_TX 64     // conv tile width  (threads along x)
#define LENSED_CONV_TY 4      // threads along y
#define LENSED_CONV_RY 8      // output rows per thread
#define LENSED_CONV_TH (LENSED_CONV_TY * LENSED_CONV_RY)  // conv tile height
#define LENSED_CONV_BS (LENSED_CONV_TX * LENSED_CONV_TY)
#define LENSED_PSF_MAX 63     // largest PSF side the conv tile is sized for

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
    int quad_off, nq, pad0, pad1;        // node offset into quad[] (float4 units), node count
};
