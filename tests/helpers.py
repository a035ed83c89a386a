"""Shared helpers for the parity tests: move a seeded scene to the GPU, run the
product path, and read the private buffers through the published layout."""
import numpy as np
import torch

import gsparc_hip as gh
import wireless_rasterizer as wr


def to_dev(scene, dev="cuda"):
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    return dict(means=t(scene.means), scales=t(scene.scales), rots=t(scene.rotations), opac=t(scene.opacities),
                mlp=t(scene.mlp), tx=t(scene.tx_pos), rx=t(scene.rx_pos), views=t(scene.views))


def gpu_forward(scene, dev="cuda", emission=None, keep_keys=True):
    d = to_dev(scene, dev)
    out = wr.forward_batched(d["means"], d["mlp"], d["opac"], d["scales"], d["rots"], scene.H, scene.W, 1.0,
                             d["views"], d["tx"], d["rx"], emission=emission, keep_keys=keep_keys)
    torch.cuda.synchronize()
    return d, out


def view_f32(buf, off, n):
    return buf[off:off + 4 * n].view(torch.float32)


def view_u32(buf, off, n):
    return buf[off:off + 4 * n].view(torch.int32)


def read_state(scene, out):
    """Private state of a forward: per-(receiver, Gaussian) records and the
    binning / image arrays, as numpy."""
    K, ore, oim, radii, geom, binbuf, img = out
    P, R, W, H = scene.P, scene.R, scene.W, scene.H
    L = gh.layout(P, R, W, H, K)
    base = lambda b: (b.data_ptr() + 255) // 256 * 256 - b.data_ptr()
    g0, i0, b0 = base(geom), base(img), base(binbuf)
    T = ((W + 15) // 16) * ((H + 15) // 16)
    st = {}
    st["splat"] = view_f32(geom, g0 + L.splat, R * P * 8).reshape(R, P, 8).cpu().numpy()
    st["sig"] = view_f32(geom, g0 + L.sig, R * P * 2).reshape(R, P, 2).cpu().numpy()
    st["depth"] = view_f32(geom, g0 + L.depth, R * P).reshape(R, P).cpu().numpy()
    st["cov3"] = view_f32(geom, g0 + L.cov3, P * 6).reshape(P, 6).cpu().numpy()
    st["final_T"] = view_f32(img, i0 + L.final_t, R * W * H).reshape(R, H, W).cpu().numpy()
    st["n_contrib"] = view_u32(img, i0 + L.n_contrib, R * W * H).reshape(R, H, W).cpu().numpy().astype(np.uint32)
    st["ranges"] = view_u32(img, i0 + L.ranges, R * T * 2).reshape(R, T, 2).cpu().numpy().astype(np.uint32)
    if K > 0:
        st["plist"] = view_u32(binbuf, b0 + L.plist, K).cpu().numpy().astype(np.uint32)
        st["keys"] = view_u32(binbuf, b0 + L.keys_sorted, K).cpu().numpy().astype(np.uint32)
    else:
        st["plist"] = np.zeros(0, np.uint32)
        st["keys"] = np.zeros(0, np.uint32)
    st["radii"] = radii.cpu().numpy()
    st["out_re"] = ore.cpu().numpy()
    st["out_im"] = oim.cpu().numpy()
    st["K"] = K
    return st
