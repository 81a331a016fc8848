"""Batch HRTEM imaging with randomly drawn aberrations, in the call shape of the reference's Python
tool ``cuda_stem64.py``: ``stem(para_part1, para_part2, myrandrange, U, V, mywave, nums)`` with
``producer(q, bat, d=0)`` putting ``(myinten [bat, N, N], newPp2 [bat, 49], bat)`` per batch and
``(0, 0, 0)`` at the end (``cuda_stem64.py:46-55,164-704``), and ``consumer(q, commpara, file_name,
file_path, position)`` writing the cropped, min-max normalised images (``:707-744``).

The arithmetic runs in the engine (``totem_hrtem_images``): FP64 transfer functions, the engine's
Stockham inverse FFT, FP64 intensities.  Differences from the reference: the uniform draws come
from a seeded numpy generator instead of curand (``:667``), and images are written as ``.npy``
arrays (the reference's tifffile is not part of this build).
"""
import os

import numpy as np

from . import engine

PP1_NAMES = ("sampling", "mtotal", "alafa", "lambda", "gmax", "gfsf", "delta_yita", "tilt", "phitilt", "tiltx",
             "tilty")
PP2_NAMES = ("lambda", "focus", "A1", "phiA1", "A2", "phiA2", "B2", "phiB2", "Cs", "A3", "phiA3", "S3", "phiS3",
             "A4", "phiA4", "D4", "phiD4", "B4", "phiB4", "A5", "phiA5", "C5", "S5", "phiS5", "D5", "phiD5",
             "A1x", "A1y", "A2x", "A2y", "B2x", "B2y", "A3x", "A3y", "S3x", "S3y", "C3", "A4x", "A4y", "D4x", "D4y",
             "B4x", "B4y", "A5x", "A5y", "S5x", "S5y", "D5x", "D5y")
MRR_NAMES = PP2_NAMES[1:22]


def images(para_part1, pp2, mrr, U, V, mywave, ai):
    """(myinten [B, N, N], newPp2 [B, 49]) for the uniform draws ai [B, 21]."""
    n = np.asarray(mywave).shape[0]
    return engine.hrtem_images(n, para_part1["lambda"], para_part1["alafa"], para_part1["delta_yita"],
                               para_part1["gfsf"], int(para_part1["mtotal"]), U, V, mywave, pp2, mrr, ai)


class stem:  # noqa: N801 -- the reference's class name
    def __init__(self, para_part1, para_part2, myrandrange, U, V, mywave, nums, seed=0):
        self.para_part1 = dict(para_part1)
        self.pp2 = np.asarray(para_part2, dtype=np.float64)
        self.mrr_list = np.asarray(myrandrange, dtype=np.float64)
        self.u = np.asarray(U, dtype=np.float64)
        self.v = np.asarray(V, dtype=np.float64)
        self.mywave = np.asarray(mywave, dtype=np.complex128)
        self.NUMS = int(nums)
        self.rng = np.random.Generator(np.random.PCG64(seed))

    def producer(self, q, bat, d=0):
        """Images in batches of ``bat`` until ``nums`` are made (the last batch is full, like the
        reference's ``count = ceil(nums / bat)`` loop), then the (0, 0, 0) end marker."""
        if d:
            import torch
            torch.cuda.set_device(d)
        count = (self.NUMS + bat - 1) // bat
        for _ in range(count):
            ai = self.rng.random((bat, 21))
            myinten, new_pp2 = images(self.para_part1, self.pp2, self.mrr_list, self.u, self.v, self.mywave, ai)
            q.put((myinten, new_pp2, bat))
        q.put((0, 0, 0))


def consumer(q, commpara, file_name, file_path, position):
    """Crops ``myinten[b, position[0]:position[1], position[2]:position[3]]``, normalises it to [0, 1]
    and saves it as ``<file_path>/<file_name + k>.npy`` together with the drawn parameters; returns the
    number of images written (``cuda_stem64.py:707-744``)."""
    os.makedirs(file_path, exist_ok=True)
    k = 0
    while True:
        myinten, new_pp2, batches = q.get()
        if batches <= 0:
            return k
        for b in range(batches):
            img = myinten[b, position[0]:position[1], position[2]:position[3]]
            img = (img - img.min()) / (img.max() - img.min())
            np.save(os.path.join(file_path, "%d.npy" % (file_name + k)), img.astype(np.float32))
            np.save(os.path.join(file_path, "%d_para.npy" % (file_name + k)), new_pp2[b])
            k += 1


def load_inputs(mat):
    """The tool's inputs from a loaded ``pydata.mat`` dict (``cuda_stem64.py:747-816``): returns
    (para_part1 dict, para_part2 [49], myrandrange [21], U, V, mywave, othervar, commpara)."""
    p1 = mat["para_part1"][0][0]
    para_part1 = {name: float(np.asarray(p1[i]).ravel()[0]) for i, name in enumerate(PP1_NAMES)}
    pp2 = np.array([float(np.asarray(x).ravel()[0]) for x in mat["para_part2"][0][0]])
    mrr = np.array([float(np.asarray(x).ravel()[0]) for x in mat["myrandrange"][0][0]])
    return (para_part1, pp2, mrr, np.asarray(mat["U"], dtype=np.float64), np.asarray(mat["V"], dtype=np.float64),
            np.asarray(mat["mywave"], dtype=np.complex128), np.asarray(mat["othervar"]).ravel(),
            np.asarray(mat["commpara"]).ravel())
