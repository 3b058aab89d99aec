/*
 * trlnet.h -- learned start vector for trlan(): a 4-layer 3x3 convolutional
 * net (1 -> 32 -> 32 -> 32 -> 1 channels, ReLU, periodic padding) from a
 * lattice's on-site potential to a start vector, trained and run with BF16
 * MFMA kernels (FP32 accumulation, FP32 master weights, Adam).  No
 * counterpart in the reference: its output enters trlan() as the start
 * vector evec(:, nec) with info->guess = 1 (SRC/trlcore.c:1577-1763).
 *
 * Parameters, flat FP32 (trlnet_param_count() values): W1[co][t] (32x9),
 * b1[32], W2[co][t][ci] (32x9x32), b2[32], W3 (as W2), b3[32],
 * W4[t][ci] (9x32), b4[1]; tap t = 3*(dy+1) + (dx+1).
 * Images: nimg images of h x w, row-major, one FP32 value per site.
 * Loss per image: 1 - (o.t)^2 / (|o|^2 |t|^2) (sign- and scale-free),
 * averaged over the images.
 */
#ifndef NUTRLAN_AMD_TRLNET_H
#define NUTRLAN_AMD_TRLNET_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TRLNET_EINVAL = -1,
    TRLNET_ENOMEM = -2,
    TRLNET_EHIP = -3,
    TRLNET_ENODEV = -4
};

typedef struct trlnet trlnet;

int trlnet_param_count(void);
/* He-normal initial weights from `seed`, zero biases. */
int trlnet_create(trlnet **net, unsigned long long seed);
int trlnet_destroy(trlnet *net);
int trlnet_get_params(trlnet *net, float *out, int n);
int trlnet_set_params(trlnet *net, const float *in, int n);
/* out = net(in) for nimg images of h x w. */
int trlnet_forward(trlnet *net, int nimg, int h, int w, const float *in, float *out);

/* trlnet_forward over groups of images of different sizes in one call:
 * group g holds nimg[g] images of h[g] x w[g]; in/out hold the groups back
 * to back (sum of nimg*h*w floats).  One upload, one download and one
 * device synchronisation for all groups. */
int trlnet_forward_many(trlnet *net, int ngroups, const int *nimg, const int *h, const int *w,
                        const float *in, float *out);

/* trlnet_forward_many with the outputs left on the device: *out_device is
 * the device address of the groups' outputs (same layout, FP32), valid
 * until the next call on this net or its destruction.  For start vectors
 * that go straight into a batch (trl_batch_job.dstart). */
int trlnet_forward_many_device(trlnet *net, int ngroups, const int *nimg, const int *h,
                               const int *w, const float *in, const float **out_device);

/* trlnet_forward_many_device with the inputs taken from FP64 images, one
 * pointer per image (groups back to back, nimg[g] images of h[g] x w[g]
 * each): the library forms fl32(img * scale) in its pinned staging buffer
 * on several host threads, so the caller does no conversion or copy. */
int trlnet_forward_many_device_f64(trlnet *net, int ngroups, const int *nimg, const int *h,
                                   const int *w, const double *const *img, double scale,
                                   const float **out_device);

/* Loss and its gradient w.r.t. the parameters (no update). */
int trlnet_loss_grad(trlnet *net, int nimg, int h, int w, const float *in,
                     const float *target, float *loss, float *grad);
/* One Adam step (beta 0.9 / 0.999) of learning rate lr; *loss before it. */
int trlnet_train_step(trlnet *net, int nimg, int h, int w, const float *in,
                      const float *target, float lr, float *loss);

#ifdef __cplusplus
}
#endif
#endif
