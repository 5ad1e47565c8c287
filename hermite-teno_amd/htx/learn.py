"""The learned troubled-face classifier (include/htx.h, "Learned troubled-face
classifier"): training set, GPU training and evaluation.

A face's label is 1 when plain FP64 TENO5 keeps every candidate stencil on
both sides for all four characteristic fields; the classifier (MLP
12 -> 64 -> 64 -> 1, GELU) predicts it from normalised density and pressure
differences and TENO smoothness contrasts of the face's 6-cell stencil.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import CLS_NPARAM, check, fptr

CLS_IN, CLS_H = 12, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + CLS_IN * CLS_H
OFF_W2 = OFF_B1 + CLS_H
OFF_B2 = OFF_W2 + CLS_H * CLS_H
OFF_W3 = OFF_B2 + CLS_H
OFF_B3 = OFF_W3 + CLS_H


def unpack(params):
    """(W1 [12, 64], b1, W2 [64, 64], b2, w3, b3) views of a parameter vector."""
    p = np.asarray(params, dtype=np.float32)
    return (p[OFF_W1:OFF_B1].reshape(CLS_IN, CLS_H), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(CLS_H, CLS_H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


class Trainer:
    """Device-resident training set plus the FP32 parameters being trained."""

    def __init__(self, capacity):
        lib = _lib.load()
        h = ctypes.c_void_p()
        check(lib.htx_trainer_create(int(capacity), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().htx_trainer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_state(self, solver):
        """Faces (both families) of the solver's current state."""
        check(_lib.load().htx_trainer_add_state(self._h, solver._h))

    def size(self):
        n, ns = ctypes.c_longlong(), ctypes.c_longlong()
        check(_lib.load().htx_trainer_size(self._h, ctypes.byref(n), ctypes.byref(ns)))
        return n.value, ns.value

    def read(self, start, n):
        feat = np.empty((n, CLS_IN), dtype=np.float32)
        lab = np.empty(n, dtype=np.uint8)
        check(_lib.load().htx_trainer_read(self._h, int(start), int(n), fptr(feat),
                                           lab.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))))
        return feat, lab

    def init(self, seed=0):
        check(_lib.load().htx_trainer_init(self._h, int(seed)))

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == CLS_NPARAM
        check(_lib.load().htx_trainer_set_params(self._h, fptr(p)))

    def params(self):
        p = np.empty(CLS_NPARAM, dtype=np.float32)
        check(_lib.load().htx_trainer_get_params(self._h, fptr(p)))
        return p

    def train(self, epochs, batch=32768, lr=3e-3, troubled_weight=4.0, seed=1):
        """Adam epochs over the set; returns the mean minibatch loss per epoch."""
        loss = np.zeros(max(epochs, 1), dtype=np.float64)
        check(_lib.load().htx_trainer_train(self._h, int(epochs), int(batch), float(lr), float(troubled_weight),
                                            int(seed), loss.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return loss[:epochs]

    def gradient(self, troubled_weight=4.0):
        """Full-set gradient of the training loss at the current parameters
        and the loss."""
        g = np.empty(CLS_NPARAM, dtype=np.float32)
        l = ctypes.c_double()
        check(_lib.load().htx_trainer_gradient(self._h, float(troubled_weight), fptr(g), ctypes.byref(l)))
        return g, l.value

    def evaluate(self, threshold=0.0, want_logits=False, inference_rounding=False):
        """FP32 forward over the set: confusion counts at the logit threshold
        (and the logits); with inference_rounding, rounded to bf16 where the
        solver's matrix-core inference rounds."""
        c = (ctypes.c_longlong * 4)()
        n, _ = self.size()
        lg = np.empty(n, dtype=np.float32) if want_logits else None
        fn = _lib.load().htx_trainer_eval_inference if inference_rounding else _lib.load().htx_trainer_eval
        check(fn(self._h, float(threshold), c, fptr(lg) if lg is not None else None))
        out = {"smooth_as_smooth": c[0], "troubled_as_smooth": c[1], "troubled_as_troubled": c[2],
               "smooth_as_troubled": c[3]}
        if want_logits:
            out["logits"] = lg
        return out


DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def default_classifier():
    """(parameters, logit threshold) of the classifier shipped with the
    package (htx/data/classifier_teno5.*, written by htx.train_classifier)."""
    import json

    p = np.load(os.path.join(DATA, "classifier_teno5.npy"), allow_pickle=False)
    with open(os.path.join(DATA, "classifier_teno5.json")) as f:
        thr = float(json.load(f)["threshold"])
    return p.astype(np.float32), thr
