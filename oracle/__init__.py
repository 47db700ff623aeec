"""TEST INFRASTRUCTURE ONLY — CPU oracle for the SRF projector hot path.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or the timed CPU
baseline).  The product package ``srf_amd`` never imports it.

Contents
  oracle.lib         ctypes binding of liboracle.so (C restatement of
                     tor.cu.cc / siddon.cu.cc in FP64 and FP32)
  oracle.preprocess  numpy restatement of preprocess/function/on_tor_lors.py
  oracle.effmap      numpy restatement of preprocess/merge_map.py and the
                     initial-image crop of data/loader.py
  oracle.recon       MLEM / OSEM driver semantics (model/recon_step.py,
                     graph/reconstruction/{worker,master}.py)
"""
