"""MI355X-native Lucas-Lehmer engine: a drop-in for the reference's ``lucaslehmer`` module.

Native parts (built in-tree by ``build.py``):
  libllhip.so        HIP kernels for gfx950 + the C ABI of include/llhip.h
  lucaslehmer*.so    pybind11 module with the reference's API (import name ``lucaslehmer``)
  lucaslehmer_cuda*.so batch module with the reference's GPU module name and API

``add_dropin_path()`` puts this directory on ``sys.path`` so that ``import lucaslehmer``
resolves to the drop-in, exactly as the reference's module is imported.
"""
import os
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))


def add_dropin_path() -> None:
    if PKG_DIR not in sys.path:
        sys.path.insert(0, PKG_DIR)
