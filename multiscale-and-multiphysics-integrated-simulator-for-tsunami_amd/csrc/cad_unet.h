// cad_unet.h -- geometry of the U-Net's volumes, shared by the network
// (cad_unet.hip) and the learned preconditioner that feeds it (cad_learned.hip).
#pragma once

// One level's batch of nb boxes of nx x ny x nz voxels, each stored with a
// one-voxel zero halo: padded sizes px = nx+2, py = ny+2, pz = nz+2.
struct UGeo {
    int nx, ny, nz, nb;
    int px, py, pz;
};

typedef struct cad_unet cad_unet;
int cad_unet_forward_impl(cad_unet *u);

extern "C" {
hipStream_t cad_unet_stream(cad_unet *u);
UGeo cad_unet_geo(cad_unet *u);
int cad_unet_param_count(void);
double cad_unet_mac_per_voxel(void);
int cad_unet_create(int nx, int ny, int nz, int nb, int train, void *stream, cad_unet **out);
void cad_unet_destroy(cad_unet *u);
int cad_unet_set_params(cad_unet *u, const float *host);
int cad_unet_get_params(cad_unet *u, float *host);
int cad_unet_init_params(int seed, float *host);
void *cad_unet_input(cad_unet *u);
float *cad_unet_output(cad_unet *u);
int cad_unet_forward(cad_unet *u);
int cad_unet_set_depth(cad_unet *u, int nz);
int cad_unet_set_input(cad_unet *u, const void *dev);
int cad_unet_get_output(cad_unet *u, float *dev);
int cad_unet_train_step(cad_unet *u, const float *target, const float *mask, float lr, float *loss);
int cad_unet_train_step_dy(cad_unet *u, const float *dy, float lr);
int cad_unet_gradients_dy(cad_unet *u, const float *dy, float *host_grad);
int cad_unet_gradients(cad_unet *u, const float *target, const float *mask, float *host_grad,
                       float *loss);
}
