// Host interface of the BF16 U-Net (tl_unet.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tl {

struct UNet;

UNet* unet_create(uint64_t seed);
void unet_destroy(UNet* u);
// Forward pass on device planes: r (scaled residual) and f (density
// feature), n x h x w pixels (h, w multiples of 8); output channel 0 into z.
void unet_apply(UNet& u, const float* r, const float* f, float* z, int n, int h, int w);
// Loss of output channel 0 against target and the gradients of every
// parameter; returns the loss.  With kx, ky (per-sample face coefficients,
// kx the west face and ky the south face of each pixel) the loss is the
// A-norm sum(d.A d) / sum(t.A t) of d = out - target over the batch, A the
// 5-point operator with no flux across the sample's edge (after smoothing:
// see unet_set_loss_smoothing); without them the mean squared error.
float unet_loss_grad(UNet& u, const float* r, const float* f, const float* target, const float* kx, const float* ky,
                     int n, int h, int w);
void unet_adam(UNet& u, float lr);
// The A-norm loss measures the error after `sweeps` weighted Jacobi sweeps
// (weight omega) of the output, as the flexible CG applies them after the
// network: per sample sum(Md.A Md) / max(sum(Mt.A Mt), batch mean), averaged,
// M = (I - omega D^-1 A)^sweeps.
// sweeps = 0 gives the plain A-norm loss.  Default 4 sweeps, omega 0.8.
void unet_set_loss_smoothing(UNet& u, int sweeps, float omega);
void unet_set_stream(UNet& u, hipStream_t s);  // the stream all kernels of u run on
hipStream_t unet_stream(UNet& u);
long unet_param_count(const UNet& u);
// Multiply-accumulates of one forward pass over n x h x w pixels.
double unet_forward_macs(int n, int h, int w);
double unet_forward_bytes(int n, int h, int w);  // HBM bytes of one forward pass, each tensor once
// Whether the 3x3 convolutions stage their input tile in LDS (the default)
// or read it through L2 per tap; both give bitwise the same output.
void set_conv_staging(bool on);
// Parameters and gradients in layer order, each layer W[cout][9][cin] then b[cout].
void unet_get_params(const UNet& u, float* host);
void unet_set_params(UNet& u, const float* host);
void unet_get_grads(const UNet& u, float* host);

}  // namespace tl
