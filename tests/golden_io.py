"""Loading helpers for the reference-written fixtures in tests/golden/."""
import hashlib
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_manifest = None


def manifest():
    global _manifest
    if _manifest is None:
        with open(os.path.join(GOLDEN, "manifest.json")) as f:
            _manifest = json.load(f)
    return _manifest


def load(case):
    with np.load(os.path.join(GOLDEN, case + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def digest(case, name):
    return manifest()[case]["arrays"][name]["sha256"]


def meta(case):
    return manifest()[case]["meta"]
