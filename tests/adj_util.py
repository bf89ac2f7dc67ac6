"""Helpers for the adjoint tests: perturbed copies of a generated scene
description, the linear Full image of an oracle render, small argument sets."""
from __future__ import annotations

import ctypes as C

import numpy as np

import raym0nade_amd as rm
from raym0nade_amd._lib import RmMaterial, RmSceneDesc, RmTexture


def perturbed_desc(gen, dtex=None, drough=None, demis=None):
    """A copy of gen.desc with texel / roughness / emission offsets; returns
    (desc, keepalive)."""
    d = gen.desc
    nm, nt = d.n_materials, d.n_textures
    mats = (RmMaterial * nm)()
    for i in range(nm):
        mats[i] = RmMaterial.from_buffer_copy(d.materials[i])
        if drough is not None:
            mats[i].roughness = np.float32(mats[i].roughness + drough[i])
        if demis is not None:
            for c in range(3):
                mats[i].emission[c] = np.float32(mats[i].emission[c] + demis[i, c])
    texs = (RmTexture * max(nt, 1))()
    keep = [mats, texs]
    off = 0
    for t in range(nt):
        T = d.textures[t]
        n = T.width * T.height * 3
        arr = np.ctypeslib.as_array(T.rgb, shape=(n,)).copy()
        if dtex is not None:
            arr = (arr + dtex[off:off + n]).astype(np.float32)
        off += n
        keep.append(arr)
        texs[t].width, texs[t].height = T.width, T.height
        texs[t].rgb = arr.ctypes.data_as(C.POINTER(C.c_float))
    nd = RmSceneDesc.from_buffer_copy(d)
    nd.materials = C.cast(mats, C.POINTER(RmMaterial))
    nd.textures = C.cast(texs, C.POINTER(RmTexture))
    return nd, keep


def texel_count(gen) -> int:
    d = gen.desc
    return sum(d.textures[t].width * d.textures[t].height * 3 for t in range(d.n_textures))


def linear_image(gbuf, rad, exposure):
    """Photo::shade(Full) without the tone curve from oracle-layout buffers."""
    return (gbuf[..., 12:15] * (rad[0, ..., :3] + rad[2, ..., :3]) + rad[1, ..., :3] + rad[3, ..., :3] +
            gbuf[..., 9:12] * exposure).astype(np.float64)


def small_args(gen, width, spp, seed=1):
    a = rm.RmRenderArgs.from_buffer_copy(gen.args)
    a.accuracy = np.float32(a.accuracy * a.width / width)
    a.width = a.height = width
    a.spp = spp
    a.seed = seed
    return a
