"""Unitree G1 model constants (float64 numpy arrays) from g1_model.json.

The JSON is written by gen_model.py from the reference MJCF; the same numbers
are compiled into the HIP kernels through csrc/g1_model_data.h.
"""
import json
import os
from types import SimpleNamespace

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

NB, NQ, NV, NU, NC = 30, 36, 35, 29, 8


def load():
    with open(os.path.join(_HERE, "g1_model.json")) as f:
        raw = json.load(f)
    d = {k: (np.asarray(v, dtype=np.float64) if isinstance(v, list) and k != "names" else v)
         for k, v in raw.items()}
    d["parent"] = np.asarray(raw["parent"], dtype=np.int64)
    d["contact_body"] = np.asarray(raw["contact_body"], dtype=np.int64)
    d["jnt_lo"] = d["jnt_range"][0::2].copy()
    d["jnt_hi"] = d["jnt_range"][1::2].copy()
    d["total_mass"] = float(np.sum(d["mass"]))
    return SimpleNamespace(**d)


MODEL = load()
