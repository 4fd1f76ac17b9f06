"""Training of the learned con2prim bracket on the device (include/phx_mlp.h).

Pairs come from the plain solver itself: random primitive states (the batched con2prim
workload's distribution) -> prim_to_con -> plain con2prim -> (features, root mu).  The
network trains with the library's own MFMA forward/backward and Adam kernels on shuffled
mini-batches that stay resident in device memory; the learning rate follows a cosine
schedule.

The loss is e1^2 + lam e2^2: e1 the error of the bracket centre (log mu z0), e2 that of the
log half-width, whose target is the log of the centre's own relative error.  That target
is noise of order one however well the centre is fitted, so with a weight near one its
gradient swamps the centre's in the layers the two outputs share, and Adam's
per-parameter scaling then sizes the shared layers' steps by that noise: lam = 0.05 left
the centre at a median relative error of 4.1e-3 on the batched workload, lam = 1e-5 at
3.8e-4 (200,000 steps).  The width's own output row is scaled by Adam alone, so it still
trains at the full rate.  width_scale multiplies the predicted half-width: 4 / 5 / 6 miss
the root in 4.4 / 1.5 / 0.7 % of the states.  That default suits a population whose
centre errors are alike; on torus cells (floored atmosphere and disc, errors decades
apart) the half-width needs the shared layers shaped by its own loss: lam = 0.05 and
width_scale 4 miss 4 % of a torus's cold cells, lam = 1e-5 and 5 miss 67 %.
"""
import numpy as np

from .device import DeviceArray, synchronize
from .workloads import batched_con2prim_states


def solver_pairs(phx, p, rng, n):
    """n states of the batched workload and the plain solver's results on them."""
    coords, prim = batched_con2prim_states(rng, n)
    d_coords = DeviceArray.from_host(coords)
    d_prim = DeviceArray.from_host(prim)
    d_cons = DeviceArray((8, n))
    phx.prim_to_con_batch(p, n, d_coords, d_prim, d_cons)
    del d_prim
    cons0 = d_cons.get()
    d_out = DeviceArray((10, n))
    d_mu = DeviceArray.from_host(np.zeros(n))
    d_fail = DeviceArray((n,))
    d_it = DeviceArray((n,), np.int32)
    phx.con2prim_batch(p, n, d_coords, d_cons, d_out, d_mu, d_fail, d_it)
    synchronize()
    return dict(coords=coords, cons=cons0, prim=d_out.get(), mu=d_mu.get(),
                fail=d_fail.get(), iters=d_it.get(), d_coords=d_coords)


def features(net, p, data, n):
    d_cons = DeviceArray.from_host(data["cons"])
    d_f = DeviceArray((4, n), np.float32)
    net.features(p, n, data["d_coords"], d_cons, d_f)
    synchronize()
    return d_f.get()


def train_bracket(phx, p, rng, n_pairs=1 << 22, steps=6000, batch=16384, lr=2e-3, lam=1e-5,
                  width_scale=5.0, width_min=1e-7, seed=3, log=None):
    """A trained Mlp (phoebus_amd.api.Mlp) and the loss history."""
    data = solver_pairs(phx, p, rng, n_pairs)
    return train_bracket_on(phx, p, rng, data, steps=steps, batch=batch, lr=lr, lam=lam,
                            width_scale=width_scale, width_min=width_min, seed=seed, log=log)


def train_bracket_on(phx, p, rng, data, steps=6000, batch=16384, lr=2e-3, lam=1e-5,
                     width_scale=5.0, width_min=1e-7, seed=3, log=None):
    """Train on solver pairs data = dict(coords [2][n] (r, theta or X1, X2), cons [8][n],
    mu [n], fail [n], d_coords = coords on the device): the features are standardized with
    the pairs' own mean and spread."""
    from .api import Mlp

    n_pairs = data["cons"].shape[1]
    keep = data["fail"] == 1.0
    net = Mlp(phx, seed=seed)
    net.set_normalization([0, 0, 0, 0], [1, 1, 1, 1])
    raw = features(net, p, data, n_pairs)[:, keep]
    mean, std = raw.mean(axis=1), raw.std(axis=1) + 1e-6
    net.set_normalization(mean, std, width_scale, width_min)
    f = features(net, p, data, n_pairs)[:, keep]
    mu = data["mu"][keep].astype(np.float32)
    batch = min(batch, f.shape[1])
    nb = max(1, f.shape[1] // batch)
    perm = rng.permutation(f.shape[1])[:nb * batch]
    fb = np.ascontiguousarray(f[:, perm].reshape(4, nb, batch).transpose(1, 0, 2))
    d_f = DeviceArray.from_host(fb)
    d_mu = DeviceArray.from_host(mu[perm])
    d_loss = DeviceArray((1,), np.float32)
    losses = []
    for s in range(steps):
        lr_s = lr * 0.5 * (1.0 + np.cos(np.pi * s / steps)) + 1e-5
        b = s % nb
        net.train_step(batch, d_f.view(b * 4 * batch, (4, batch)),
                       d_mu.view(b * batch, (batch,)), lr_s, lam, d_loss)
        if s % 500 == 0 or s == steps - 1:
            synchronize()
            losses.append(float(d_loss.get()[0]))
            if log:
                log(f"step {s} loss {losses[-1]:.5f}")
    synchronize()
    net.norm = (mean, std)
    return net, losses


def mesh_pairs(mesh, gids, stride=1):
    """Solver pairs from the interior cells of the blocks gids of a mesh (phoebus_amd.api.Mesh)
    after a step: the conserved state, its con2prim root and fail flag, and the cells'
    (X1, X2) -- the metric of the torus meshes does not depend on X3."""
    from . import _abi as A

    out = {"coords": [], "cons": [], "mu": [], "fail": []}
    for g in gids:
        b = mesh.blocks[g]
        ng = b.ng
        I = (slice(ng, ng + b.nx[2]), slice(ng, ng + b.nx[1]), slice(ng, ng + b.nx[0]))
        cons = mesh.get(g, A.DRIVER_CONS)[(slice(None),) + I].reshape(8, -1)[:, ::stride]
        mu = mesh.get(g, A.DRIVER_C2P_MU)[I].reshape(-1)[::stride]
        fail = mesh.get(g, A.DRIVER_FAIL)[I].reshape(-1)[::stride]
        x1 = b.xmin[0] + (np.arange(b.nx[0]) + 0.5) * b.dx[0]
        x2 = b.xmin[1] + (np.arange(b.nx[1]) + 0.5) * b.dx[1]
        X1 = np.broadcast_to(x1[None, None, :], (b.nx[2], b.nx[1], b.nx[0])).reshape(-1)[::stride]
        X2 = np.broadcast_to(x2[None, :, None], (b.nx[2], b.nx[1], b.nx[0])).reshape(-1)[::stride]
        out["coords"].append(np.stack([X1, X2]))
        out["cons"].append(cons)
        out["mu"].append(mu)
        out["fail"].append(fail)
    return {k: np.concatenate(v, axis=-1) for k, v in out.items()}
