/*
 * cffc_unet.h -- C ABI of the learned warm start for steady 2D Euler solves.
 *
 * No counterpart in the reference: CFFC starts steady solves from the
 * uniform initial state (ICs, Euler2DQuadSingleBlock.cc:1689).  This module
 * adds a 4-level convolutional U-Net that predicts the initial state of the
 * bump-channel solve from the geometry and the inflow Mach number; the plain
 * solver then converges from that state to the same residual tolerance.
 *
 * The network (about 0.49 M parameters) runs entirely on the GPU with
 * hand-written BF16 MFMA kernels (v_mfma_f32_16x16x32_bf16) and FP32
 * accumulation: forward, backward (input and weight gradients) and Adam.
 * Layout: NHWC images of H x W pixels; inputs have CFFC_UNET_IN channels,
 * outputs CFFC_UNET_OUT channels.  Weights of layer l are FP32 masters
 * [Cout][9*Cin] with k = tap*Cin + ci (tap = 3*dy + dx), biases [Cout].
 *
 * Return codes: 0 = success, < 0 = CFFC_E_* (cffc_euler2d.h).
 */
#ifndef CFFC_UNET_H
#define CFFC_UNET_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CFFC_UNET_IN 8
#define CFFC_UNET_OUT 4

typedef struct cffc_unet cffc_unet;

/* Network for batches of up to max_batch images of H x W pixels (H and W
   divisible by 8). Weights are initialised from seed (He normal). */
int  cffc_unet_create(int H, int W, int max_batch, uint64_t seed, cffc_unet **out);
void cffc_unet_destroy(cffc_unet *net);

/* Number of conv layers and the shape of layer l: {CinA, CinB, Cout, level}. */
int  cffc_unet_num_layers(const cffc_unet *net);
int  cffc_unet_layer_shape(const cffc_unet *net, int l, int shape[4]);
/* Host copies of the FP32 master weights / biases of layer l. */
int  cffc_unet_get_weights(cffc_unet *net, int l, float *w, float *b);
int  cffc_unet_set_weights(cffc_unet *net, int l, const float *w, const float *b);

/* Training data kept on the device: n samples, inputs [n][H][W][IN] (FP32,
   stored as BF16) and targets [n][H][W][OUT] (FP32). */
int  cffc_unet_set_dataset(cffc_unet *net, int n, const float *inputs, const float *targets);

/* One Adam step on the samples idx[0..batch) of the dataset; *loss receives
   the batch MSE before the step.  lr is the step size. */
int  cffc_unet_train_step(cffc_unet *net, const int32_t *idx, int batch, float lr, float *loss);

/* Forward and backward of the loss on the given samples without updating:
   loss, and the weight/bias gradients of layer l copied to gw/gb if not NULL
   (for gradient checks). */
int  cffc_unet_loss_grad(cffc_unet *net, const int32_t *idx, int batch, float *loss, int l, float *gw,
                         float *gb);

/* Predict: inputs [batch][H][W][IN] host FP32 -> outputs [batch][H][W][OUT]. */
int  cffc_unet_predict(cffc_unet *net, const float *inputs, int batch, float *outputs);

/* Check of the MFMA operand layout the kernels assume: D = A*B for a known
   16x32 and 32x16 pair computed on the device; out receives 64 lanes x 4
   values (lane-major).  Returns 0 when the launch succeeded. */
int  cffc_unet_mfma_layout_probe(float out[256]);

/* HIP stream of the network (hipStream_t). */
void *cffc_unet_stream(cffc_unet *net);

#ifdef __cplusplus
}
#endif

#endif /* CFFC_UNET_H */
