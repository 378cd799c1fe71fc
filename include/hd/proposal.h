/* hd/proposal.h -- the learned proposal that gates the full-resolution search.
 *
 * No entry point here has a counterpart in the reference, which searches
 * every DM trial of every gulp.  The proposal looks at a coarse DM-time plane
 * (the gulp dedispersed at 16x the sampling time on every 8th trial of the
 * plan, baseline-removed and RMS-normalised by the search's own stages) in
 * 64x64 bfloat16 tiles of two channels -- the plane, and the plane averaged
 * over 4 more samples (a 4x longer window around the same centre) -- and a
 * small CNN scores each tile.  A gated search
 * then runs dedispersion, normalisation, boxcars, giants and clustering only
 * on the DM units (HD_PROPOSAL_UNIT consecutive trials) whose tiles score
 * above a threshold; every value it computes is the one the full search
 * computes, so the gated candidates are full-search candidates.
 *
 * Tile (u, v) of a gulp covers coarse rows [16u-24, 16u+40); channel 0
 * holds coarse samples [32v-16, 32v+48), channel 1 the wide plane's samples
 * [8v-28, 8v+36) (wide sample j = (z[4j]+z[4j+1]+z[4j+2]+z[4j+3])/2 of the
 * coarse plane z); zero outside the planes.  Tiles are stored [64][64][2].
 * Its rows [24, 40) are unit u (trials [128u, 128u+128)); its channel-0
 * samples span [512v - 256, 512v + 768) of the gulp.  A training label marks
 * a tile whose unit and span hold the peak of a full-search candidate.
 *
 * Network: four 3x3 / stride 2 / pad 1 convolutions with 32 channels and
 * ReLU (64 -> 32 -> 16 -> 8 -> 4 pixels; 2 input channels), fc 512 -> 64
 * with ReLU, fc 64 -> 1 logit.  Matrix stages run on gfx950 MFMA with bfloat16 operands and float32
 * accumulation; activations are stored as bfloat16, the master weights and
 * the Adam moments as float32.  The flat float32 parameter layout is
 *   w1[32][9][2] b1[32] w2[32][9][32] b2[32] w3 b3 w4 b4
 *   fc1_w[64][512] fc1_b[64] fc2_w[64] fc2_b[1]
 * (conv weights [out][ky*3+kx][in]; fc1 input index (y*4+x)*32 + channel).
 */
#pragma once

#include "hd/types.h"
#include "hd/error.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HD_PROPOSAL_TILE 64        /* tile side (coarse rows and samples) */
#define HD_PROPOSAL_TIME 16        /* time decimation of the coarse plane */
#define HD_PROPOSAL_DM 8           /* trials per coarse row */
#define HD_PROPOSAL_UNIT 128       /* trials per gated unit (16 coarse rows) */
#define HD_PROPOSAL_TSTEP 32       /* coarse samples between tile centres */

typedef struct hd_cnn_t* hd_cnn;

/* Number of float32 parameters (61249). */
hd_size hd_cnn_param_count(void);

/* A network on device gpu_id for batches of at most max_batch tiles, with
 * its own stream, parameters, gradients, Adam moments and activations. */
hd_error hd_cnn_create(hd_cnn* net, int gpu_id, hd_size max_batch);
void hd_cnn_destroy(hd_cnn net);

/* He-normal weights from a seeded generator, zero biases, zero moments. */
hd_error hd_cnn_init_params(hd_cnn net, unsigned long long seed);
hd_error hd_cnn_set_params(hd_cnn net, const hd_float* host_params);
hd_error hd_cnn_get_params(hd_cnn net, hd_float* host_params);

/* Logits of n tiles (device, bfloat16 bits [n][64][64][2]); d_logits (device,
 * n floats) may be NULL, then hd_cnn_get_logits reads them.  The
 * activations are kept for hd_cnn_backward. */
hd_error hd_cnn_forward(hd_cnn net, const unsigned short* d_tiles, hd_size n, hd_float* d_logits);
hd_error hd_cnn_get_logits(hd_cnn net, hd_float* host, hd_size n);

/* Gradients of the weighted mean binary cross-entropy of the last forward
 * batch (labels 0/1 and weights in device memory, n floats each).  loss
 * (host, may be NULL) receives the loss. */
hd_error hd_cnn_backward(hd_cnn net, const hd_float* d_labels, const hd_float* d_weights,
                         hd_float* loss);
hd_error hd_cnn_get_grads(hd_cnn net, hd_float* host_grads);

/* One Adam step (decoupled weight decay) on the gradients of the last
 * backward; refreshes the bfloat16 weight copies. */
hd_error hd_cnn_adam(hd_cnn net, hd_float lr, hd_float beta1, hd_float beta2, hd_float eps,
                     hd_float weight_decay);

/* ---- coarse plane and gating on a pipeline ---- */

/* Sets the network (host float32 parameters, hd_cnn_param_count values) and
 * the logit threshold a tile must exceed for its unit to be searched.  With
 * params NULL gating is switched off.  While it is on, hd_execute_batch and
 * hd_execute_device search only the proposed units of each gulp. */
hd_error hd_set_proposal(hd_pipeline pl, const hd_float* params, hd_float threshold);

/* Computes the coarse plane of one single-beam gulp in device memory (the
 * pipeline keeps it) and returns its unit and tile-position counts. */
hd_error hd_proposal_coarse(hd_pipeline pl, const hd_byte* d_filterbank, hd_size nsamps,
                            hd_size nbits, hd_size* n_units, hd_size* n_tpos);

/* Extracts tiles (u, v) = (pos[2i], pos[2i+1]) of the last coarse plane into
 * d_tiles (device, bfloat16 bits [n][64][64][2]). */
hd_error hd_proposal_extract(hd_pipeline pl, const unsigned* pos, hd_size n,
                             unsigned short* d_tiles);

/* Copies channel `channel` (0: coarse plane, 1: wide plane) of the last
 * coarse plane of the first gulp (bfloat16 bits, rows x cols) to the host;
 * returns rows*cols, or 0 if none. */
hd_size hd_proposal_get_plane(hd_pipeline pl, int channel, unsigned short* host, hd_size cap,
                              hd_size* rows, hd_size* cols);

/* Logits of every tile of the last coarse plane with the network set by
 * hd_set_proposal (host, n_units*n_tpos floats, u major); returns the count. */
hd_size hd_proposal_scores(hd_pipeline pl, hd_float* host, hd_size cap);

/* Per gulp of the last gated call, the units that were searched: flags[g *
 * units_per_gulp + u]; returns the number written. */
hd_size hd_get_last_proposal(hd_pipeline pl, unsigned char* flags, hd_size cap,
                             hd_size* units_per_gulp);

#ifdef __cplusplus
}
#endif
