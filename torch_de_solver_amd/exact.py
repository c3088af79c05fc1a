"""Float64 reference solutions of the standard workloads, for relative-L2
error reporting.

* Legendre n = 3: P3(x) = (5x^3 - 3x)/2.
* wave u_tt = 4 u_xx, u(x,0) = sin(pi x), u_t(x,0) = 0, u(0)=u(1)=0:
  sin(pi x) cos(2 pi t).
* KdV soliton of speed 1 on a 20-periodic line: 0.5 sech^2(0.5 (x - t - 10)),
  summed over periodic images.
* Burgers u_t + u u_x = nu u_xx on [-1, 1], u(x,0) = -sin(pi x),
  u(+-1,t) = 0: exact through the Cole-Hopf transform u = -2 nu phi_x/phi,
  phi_t = nu phi_xx, phi(x,0) = exp(-k cos(pi x)), k = 1/(2 pi nu).  phi(.,0)
  is even and 2-periodic, so the whole-line heat kernel keeps phi_x(+-1) = 0
  and
      u(x,t) = int (eta/t) g(x - eta) G(eta) d eta / int g(x - eta) G(eta) d eta
  with g = phi(.,0) and G(eta) = exp(-eta^2/(4 nu t)).  Both integrals are
  evaluated with the trapezoid rule on eta = s z, s = sqrt(4 nu t), with
  log-sum-exp scaling.  The exponent -k cos(pi(x - eta)) - z^2 is at most
  k - z^2 and its maximum is at least -k, so |z| <= sqrt(2k + 40) holds all
  terms within e^-40 of the largest; the integrand is smooth and negligible
  at the window's ends, so the rule converges spectrally once the nodes
  resolve the width 1/(pi sqrt(k)) of the peaks of g.
"""

import math

import numpy as np


def relative_l2(u: np.ndarray, u_ref: np.ndarray) -> float:
    u = np.asarray(u, dtype=np.float64).reshape(-1)
    r = np.asarray(u_ref, dtype=np.float64).reshape(-1)
    return float(np.linalg.norm(u - r) / np.linalg.norm(r))


def legendre_p3(x):
    x = np.asarray(x, dtype=np.float64)
    return 0.5 * (5 * x ** 3 - 3 * x)


def wave(x, t):
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    return np.sin(np.pi * x) * np.cos(2 * np.pi * t)


def kdv_soliton(x, t, L=20.0, x0=10.0, images=2):
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    out = np.zeros(np.broadcast(x, t).shape)
    for k in range(-images, images + 1):
        out += 0.5 / np.cosh(0.5 * (x - t - x0 + k * L)) ** 2
    return out


def burgers(x, t, nu, nodes: int = None):
    """u(x, t) on arrays of equal (or broadcastable) shape."""
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    x, t = np.broadcast_arrays(x, t)
    k = 1.0 / (2 * np.pi * nu)
    xf, tf = x.reshape(-1), t.reshape(-1)
    u = np.empty_like(xf)
    zero = tf <= 0
    u[zero] = -np.sin(np.pi * xf[zero])
    idx = np.nonzero(~zero)[0]
    if idx.size:
        smax = math.sqrt(4 * nu * float(tf[idx].max()))
        zmax = max(9.0, math.sqrt(2 * k + 40))
        width = 1.0 / (np.pi * math.sqrt(k))
        m = nodes or int(min(40001, max(401, 24 * 2 * zmax * smax / width)))
        z = np.linspace(-zmax, zmax, m)            # eta = s z
        chunk = max(1, (1 << 22) // m)
        for c in range(0, idx.size, chunk):
            ii = idx[c:c + chunk]
            s = np.sqrt(4 * nu * tf[ii])[:, None]
            eta = s * z[None, :]
            a = -k * np.cos(np.pi * (xf[ii, None] - eta)) - z[None, :] ** 2
            a -= a.max(axis=1, keepdims=True)
            w = np.exp(a)
            u[ii] = np.sum(w * eta, axis=1) / np.sum(w, axis=1) / tf[ii]
    return u.reshape(x.shape)
