// zq3d_adjoint.h -- launchers of the adjoint kernels (kernels_adjoint.hip), device pointers,
// stream-ordered, templated on R = float | double.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zq3d_device.h"
#include "zq3d_kernels.h"

// longest backtrace the adjoint stores per thread (advection substeps)
#define ZQ3D_ADJ_MAX_STEPS 32

namespace zq3d {

// Force + attenuation adjoint, in place on the face adjoints ub/vb/wb and the cell adjoints Tb/Rb
// (adjoints after the stage in, before it out).  ctrl_b (nullable, 3 per cell) receives the
// adjoint of the control force.  vort: the forward's (w, |w|) of the stage's input velocity.
// gb, wbar, cb: 3N scratch each.
template <class R>
void launch_adj_force(const Dims& d, R* ub, R* vb, R* wb, R* Tb, R* Rb, R* ctrl_b, const uint8_t* occ,
                      const vec4_t<R>* vort, const ForceParams& fp, float aV, float aT, float aR, R* gb, R* wbar,
                      R* cb, hipStream_t s);

// Velocity advection adjoint, in place on ub/vb/wb.  u/v/w: the stage's input velocity;
// vel4: its cell average (modes 0 and 4).  Scratch: cellb, advb 3N; outb Nu+Nv+Nw (modes 2, 4).
template <class R>
void launch_adj_advect_velocity(const Dims& d, const R* u, const R* v, const R* w, R* ub, R* vb, R* wb,
                                const uint8_t* occ, int is_open, int mode, const AdvParams& ap,
                                const vec4_t<R>* vel4, R* cellb, R* advb, R* outb, hipStream_t s);

// Scalar advection adjoint: Tb_in/Rb_in = adjoints of the inputs Tin/Rin (overwritten), the
// velocity adjoint added to ub/vb/wb.  vel4: cell average of u/v/w (REG mode); cellb 3N scratch.
template <class R>
void launch_adj_advect_scalar(const Dims& d, const R* u, const R* v, const R* w, R* ub, R* vb, R* wb,
                              const uint8_t* occ, const R* Tin, const R* Rin, const R* Tb, const R* Rb, R* Tb_in,
                              R* Rb_in, int mac, const AdvParams& ap, const vec4_t<R>* vel4, R* cellb,
                              hipStream_t s);

// faces += transposed MAC -> cell average of the cell adjoint cb (3N SoA)
template <class R>
void launch_add_cells_to_faces(const Dims& d, const R* cb, R* ub, R* vb, R* wb, hipStream_t s);

// p[0..n) = 0 (a kernel, so it may sit in a captured graph)
template <class R>
void launch_fill_adj(R* p, size_t n, hipStream_t s);

}  // namespace zq3d
