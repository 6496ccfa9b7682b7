/* ewg_cnn.h -- learned Newton warm start for libewg: a 3D convolutional network with
 * BF16 weights and activations and FP32 accumulation on the gfx950 matrix cores.
 *
 * No counterpart in the reference: eWoms' NewtonMethod always starts from the previous
 * time level's solution (newtonmethod.hh:308-517 takes solution(0) as it is). This
 * interface adds a network that predicts the step's update from the state, and the
 * Newton driver (ewg_newton / ewg_run) takes its prediction as the initial guess when one
 * is attached; the unchanged convergence test judges the step, and a step that does not
 * converge from the prediction is redone from the plain guess.
 *
 * Network (89,122 parameters):
 *   h0 = relu(conv0(f) + b0)                 5 -> 32 channels
 *   h_l = h_{l-1} + relu(conv_l(h_{l-1}) + b_l), l = 1..3        32 -> 32
 *   y = conv4(h3) + b4                      32 -> 2 channels
 * every conv 3x3x3, stride 1, zero padding 1 (cross-correlation, as torch.nn.Conv3d).
 * Inputs f per cell: (p_w - p_ref)/p_scale, S_n, (ln K - lnk_ref)/lnk_scale,
 * (phi - phi_ref)/phi_scale, dt/dt_scale. Outputs: the update of the step,
 * dp_w = y0 * dp_scale and dS_n = y1 * ds_scale.
 *
 * Parameters are FP32 master values in one flat array, layer by layer: W_l in the order
 * [out][in][kz][ky][kx] (torch.nn.Conv3d), then b_l[out].
 *
 * Tensors passed as device pointers (void/uint16_t/float *) live in device memory of the
 * current HIP device (ewg_device_alloc gives such memory to callers without an allocator
 * of their own): features BF16 [cells][8] (channels 5..7 zero), targets and outputs FP32
 * [cells][2]. A batch stacks `batch` samples of the network's nx*ny*nz grid along z (each
 * sample zero-padded on its own). Calls are synchronous.
 */
#ifndef EWG_CNN_H
#define EWG_CNN_H

#include <stdint.h>

#include "ewg_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define EWG_CNN_NUM_PARAMS 89122

typedef struct ewg_cnn ewg_cnn;

/* Feature and target scaling; the guess takes the predicted update only where it matters:
 * dS_n where |dS_n| >= ds_min, dp_w where |dp_w| >= dp_min (0: everywhere). */
typedef struct ewg_cnn_norm {
    double p_ref, p_scale;
    double lnk_ref, lnk_scale;
    double phi_ref, phi_scale;
    double dt_scale;
    double dp_scale, ds_scale;
    double dp_min, ds_min;
} ewg_cnn_norm;

/* A network for samples of nx*ny*nz cells (nx a multiple of 16, ny, nz >= 1) and batches
 * of up to max_batch samples; parameters start at zero. */
int ewg_cnn_create(int32_t nx, int32_t ny, int32_t nz, int32_t max_batch, ewg_cnn **out);
void ewg_cnn_destroy(ewg_cnn *net);
const char *ewg_cnn_last_error(const ewg_cnn *net);

int ewg_cnn_set_params(ewg_cnn *net, const float *params);  /* host, EWG_CNN_NUM_PARAMS */
int ewg_cnn_get_params(const ewg_cnn *net, float *params);  /* host */
int ewg_cnn_set_norm(ewg_cnn *net, const ewg_cnn_norm *norm);
int ewg_cnn_get_norm(const ewg_cnn *net, ewg_cnn_norm *norm);

/* Features of a state of `ncells` cells with the network's scaling (device pointers:
 * x [cells][2] = (p_w, S_w), K [cells], phi [cells]) into feat (device, BF16 [cells][8]).
 * The state may be of any grid (training data are cut from larger grids). */
int ewg_cnn_features(ewg_cnn *net, const double *x, const double *K, const double *phi,
                     int64_t ncells, double dt, uint16_t *feat);

/* y = network(feat) for `batch` samples (device pointers). */
int ewg_cnn_forward(ewg_cnn *net, const uint16_t *feat, int32_t batch, float *y);

/* Loss 1/(2 cells) sum (y - target)^2 over a batch and its gradient with respect to the
 * parameters (host, EWG_CNN_NUM_PARAMS floats; may be NULL). feat and target are device
 * arrays of nsamples samples; idx (host, `batch` entries in 0 .. nsamples-1) picks the
 * batch's samples from them, or NULL for samples 0 .. batch-1. */
int ewg_cnn_gradients(ewg_cnn *net, const uint16_t *feat, const float *target,
                      int64_t nsamples, const int32_t *idx, int32_t batch, double *loss,
                      float *grad);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the loss above; *loss is the loss
 * before the step. */
int ewg_cnn_train_step(ewg_cnn *net, const uint16_t *feat, const float *target,
                       int64_t nsamples, const int32_t *idx, int32_t batch,
                       float learning_rate, double *loss);

/* Device time of the last ewg_cnn_forward (ms, from events around the five convolution
 * launches), for the matrix-core utilisation figure. */
int ewg_cnn_last_forward_ms(const ewg_cnn *net, double *ms);

/* Device memory of the current device for the network's data; copies are synchronous. */
int ewg_device_alloc(int64_t bytes, void **ptr);
int ewg_device_free(void *ptr);
int ewg_device_upload(void *dst, const void *src, int64_t bytes);
int ewg_device_download(void *dst, const void *src, int64_t bytes);

/* Attach a network (NULL detaches) to a context: from then on, while
 * params.warm_start is 1, every Newton solve of ewg_newton / ewg_run starts from
 * solution(1) + the network's prediction, and with storage of the previous time level
 * from solution(1) (first-iteration storage recycling presumes the plain guess); a solve
 * that fails from the prediction is redone from the plain guess. The network must have
 * been created for the context's grid with max_batch >= 1. The context does not own it.
 *
 * The prediction is evaluated on the active region only: the bounding box of the cells
 * with non-wetting saturation above 1e-8 and of the cells next to boundaries that inject
 * it, grown by 8 cells (the five 3x3x3 layers reach 5 cells, so the box's cells and 3
 * layers around them get the whole-grid prediction bit for bit); elsewhere the guess is
 * solution(1). A region of more than half the grid is evaluated on the whole grid. */
int ewg_set_warm_start(ewg_ctx *ctx, ewg_cnn *net);

/* Cells the last warm-start prediction was evaluated on (0: no active cell). */
int64_t ewg_cnn_last_inference_cells(const ewg_cnn *net);

/* Steps that started from the prediction and converged (accepted), steps redone from the
 * plain guess (rejected), the time of the predictions (s, device clock between events
 * recorded around each one on the context's stream), and the Newton error (max-norm of
 * the residual) at the first linearization of the last Newton solve, warm or plain. */
int ewg_warm_start_stats(const ewg_ctx *ctx, int64_t *accepted, int64_t *rejected,
                         double *predict_s, double *last_initial_error);

#ifdef __cplusplus
}
#endif

#endif /* EWG_CNN_H */
