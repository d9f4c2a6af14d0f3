"""Capsule membrane load on the GPU: the reference's FP64 spherical-harmonic membrane law and
its learned BF16-MFMA surrogate (libmemb.so, include/memb.h).

The FP64 law restates complagrforces / cal_spectral_force (forcing.f90:11-114,
sph.f90:1523-1755): shape coefficients, derivatives of the surface up to 4th order on the 2x
grid, neo-Hookean tension and bending loads, dealias filter.  The surrogate is an MLP from the
shape coefficients to the traction, trained and run by the library's own MFMA kernels.
Device arrays are torch tensors (torch provides device memory and streams only).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class MembError(RuntimeError):
    pass


def _ptr(t):
    return None if t is None else ctypes.c_void_p(int(t.data_ptr()))


class MembranePlan:
    """Tables for degree L and the FP64 law (memb_plan_*)."""

    def __init__(self, L: int, device: int = 0):
        self.lib = _lib.load_memb()
        h = ctypes.c_void_p()
        rc = self.lib.memb_plan_create(ctypes.byref(h), int(L), int(device))
        self.h = h
        if rc:
            msg = self.lib.memb_last_error(h).decode() if h.value else ""
            self.close()
            raise MembError(f"memb_plan_create({L}) -> {rc}: {msg}")
        self.L = int(L)
        self.npts = self.lib.memb_npts(self.h)
        self.ncoef = self.lib.memb_ncoef(self.h)

    def _check(self, rc):
        if rc:
            raise MembError(f"libmemb error {rc}: {self.lib.memb_last_error(self.h).decode()}")

    def set_stream(self, stream):
        self._check(self.lib.memb_set_stream(self.h, ctypes.c_void_p(int(stream))))

    def sync(self):
        self._check(self.lib.memb_sync(self.h))

    def traction_dev(self, x, y, z, gs, eb, fx, fy, fz, nx, ny, nz, area):
        ncap = x.numel() // self.npts
        self._check(self.lib.memb_traction_f64_dev(
            self.h, ncap, float(gs), float(eb), _ptr(x), _ptr(y), _ptr(z), _ptr(fx), _ptr(fy),
            _ptr(fz), _ptr(nx), _ptr(ny), _ptr(nz), _ptr(area)))

    def coeffs_dev(self, x, y, z, coef):
        ncap = x.numel() // self.npts
        self._check(self.lib.memb_coeffs_f64_dev(self.h, ncap, _ptr(x), _ptr(y), _ptr(z),
                                                 _ptr(coef)))

    def geometry_dev(self, x, y, z, nx, ny, nz, area, skip=None):
        """Normals and areas (memb_geometry_f64_dev); capsules with skip[c] != 0 untouched."""
        ncap = x.numel() // self.npts
        self._check(self.lib.memb_geometry_f64_dev(self.h, ncap, _ptr(x), _ptr(y), _ptr(z),
                                                   _ptr(nx), _ptr(ny), _ptr(nz), _ptr(area),
                                                   _ptr(skip)))

    def traction_select_dev(self, sel, x, y, z, gs, eb, fx, fy, fz, nx, ny, nz, area):
        """The law for the capsules with sel != 0 (int32 device tensor of ncap) only."""
        ncap = x.numel() // self.npts
        self._check(self.lib.memb_traction_select_f64_dev(
            self.h, ncap, _ptr(sel), float(gs), float(eb), _ptr(x), _ptr(y), _ptr(z), _ptr(fx),
            _ptr(fy), _ptr(fz), _ptr(nx), _ptr(ny), _ptr(nz), _ptr(area)))

    def corot_dev(self, X, x, y, z, delta, dhx, dhy, dhz, rot, amp):
        ncap = x.numel() // self.npts
        self._check(self.lib.memb_corot_f64_dev(self.h, ncap, _ptr(X), _ptr(x), _ptr(y), _ptr(z),
                                                _ptr(delta), _ptr(dhx), _ptr(dhy), _ptr(dhz),
                                                _ptr(rot), _ptr(amp)))

    def corot_apply_dev(self, flin, resid, rot, amp, fx, fy, fz):
        ncap = amp.numel()
        self._check(self.lib.memb_corot_apply_f64_dev(self.h, ncap, _ptr(flin), _ptr(resid),
                                                      _ptr(rot), _ptr(amp), _ptr(fx), _ptr(fy),
                                                      _ptr(fz)))

    def anchored_predict_dev(self, lin, m0, b, delta, delta0, rot, beta, tol, fx, fy, fz, sel,
                             drift=None, pre=None, late=None):
        """The anchored surrogate's prediction and gate (memb_anchored_predict_f64_dev)."""
        ncap = rot.numel() // 9
        self._check(self.lib.memb_anchored_predict_f64_dev(
            self.h, ncap, _ptr(lin), _ptr(m0), _ptr(b), _ptr(delta), _ptr(delta0), _ptr(rot),
            float(beta), float(tol), _ptr(fx), _ptr(fy), _ptr(fz), _ptr(sel), _ptr(drift),
            _ptr(pre), _ptr(late)))

    def anchored_gate_dev(self, delta, delta0, beta, tol, pre):
        """The drift test alone (memb_anchored_gate_f64_dev): pre[c] = 1 for rejected capsules."""
        ncap = pre.numel()
        self._check(self.lib.memb_anchored_gate_f64_dev(self.h, ncap, _ptr(delta), _ptr(delta0),
                                                        float(beta), float(tol), _ptr(pre)))

    def anchored_update_dev(self, sel, slot, rot, fx, fy, fz, lin, m_now, delta, amp, b, delta0,
                            m0, s0, refreshed):
        """New anchors for the selected capsules (memb_anchored_update_f64_dev)."""
        ncap = rot.numel() // 9
        self._check(self.lib.memb_anchored_update_f64_dev(
            self.h, ncap, _ptr(sel), _ptr(slot), _ptr(rot), _ptr(fx), _ptr(fy), _ptr(fz),
            _ptr(lin), _ptr(m_now), _ptr(delta), _ptr(amp), _ptr(b), _ptr(delta0), _ptr(m0),
            _ptr(s0), _ptr(refreshed)))

    def compact_sel_dev(self, sel, k, idx, slot):
        """The first k selected capsules (memb_compact_sel_dev)."""
        self._check(self.lib.memb_compact_sel_dev(self.h, int(sel.numel()), _ptr(sel), int(k),
                                                  _ptr(idx), _ptr(slot)))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.memb_plan_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SurrogateMLP:
    """The traction surrogate (memb_mlp_*): din -> width x nhidden (SiLU) -> dout."""

    def __init__(self, din, width, nhidden, dout, max_batch, device=0):
        self.lib = _lib.load_memb()
        h = ctypes.c_void_p()
        rc = self.lib.memb_mlp_create(ctypes.byref(h), din, width, nhidden, dout, max_batch,
                                      device)
        self.h = h
        if rc:
            msg = self.lib.memb_mlp_last_error(h).decode() if h.value else ""
            self.close()
            raise MembError(f"memb_mlp_create -> {rc}: {msg}")
        self.din, self.width, self.nhidden, self.dout, self.max_batch = (
            din, width, nhidden, dout, max_batch)
        self.nl = self.lib.memb_mlp_num_layers(self.h)

    def _check(self, rc):
        if rc:
            raise MembError(f"libmemb mlp error {rc}: {self.lib.memb_mlp_last_error(self.h).decode()}")

    def set_stream(self, stream):
        self._check(self.lib.memb_mlp_set_stream(self.h, ctypes.c_void_p(int(stream))))

    def shape(self, i):
        o, n = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.memb_mlp_layer_shape(self.h, i, ctypes.byref(o), ctypes.byref(n)))
        return o.value, n.value

    def init(self, seed):
        self._check(self.lib.memb_mlp_init(self.h, int(seed)))

    def set_norm(self, mean, scale, out_mean, out_scale):
        """x' = (x - mean) * scale; y = out_mean + out_scale * y' (scalars broadcast)."""
        f = lambda a, n: np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float32), (n,)))
        arrs = [f(mean, self.din), f(scale, self.din), f(out_mean, self.dout), f(out_scale, self.dout)]
        self._norm_arrays = arrs
        self._check(self.lib.memb_mlp_set_norm(self.h, *[a.ctypes.data_as(ctypes.c_void_p) for a in arrs]))

    def set_layer(self, i, W, b):
        W = np.ascontiguousarray(W, dtype=np.float32)
        b = np.ascontiguousarray(b, dtype=np.float32)
        self._check(self.lib.memb_mlp_set_layer(self.h, i, W.ctypes.data_as(ctypes.c_void_p),
                                                b.ctypes.data_as(ctypes.c_void_p)))

    def get_layer(self, i):
        o, n = self.shape(i)
        W = np.zeros((o, n), np.float32)
        b = np.zeros(o, np.float32)
        self._check(self.lib.memb_mlp_get_layer(self.h, i, W.ctypes.data_as(ctypes.c_void_p),
                                                b.ctypes.data_as(ctypes.c_void_p)))
        return W, b

    def get_grad(self, i):
        o, n = self.shape(i)
        W = np.zeros((o, n), np.float32)
        b = np.zeros(o, np.float32)
        self._check(self.lib.memb_mlp_get_grad(self.h, i, W.ctypes.data_as(ctypes.c_void_p),
                                               b.ctypes.data_as(ctypes.c_void_p)))
        return W, b

    def forward(self, x, y, batch=None):
        batch = x.shape[0] if batch is None else batch
        self._check(self.lib.memb_mlp_forward(self.h, int(batch), _ptr(x), _ptr(y)))

    def train_step(self, x, t, lr, wt=None, batch=None):
        batch = x.shape[0] if batch is None else batch
        loss = ctypes.c_float()
        self._check(self.lib.memb_mlp_train_step(self.h, int(batch), _ptr(x), _ptr(t), _ptr(wt),
                                                 float(lr), ctypes.byref(loss)))
        return loss.value

    def grad_only(self, x, t, wt=None, batch=None):
        batch = x.shape[0] if batch is None else batch
        loss = ctypes.c_float()
        self._check(self.lib.memb_mlp_grad_only(self.h, int(batch), _ptr(x), _ptr(t), _ptr(wt),
                                                ctypes.byref(loss)))
        return loss.value

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.memb_mlp_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
