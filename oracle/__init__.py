"""CPU oracle for the filterbank dedispersion hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, on the CPU, the reference algorithm of every function on the
path (sigpyproc ``core/kernels.py``, ``params.py``, ``header.py``, ``readers.py``,
``base.py``, ``core/filters.py``, ``core/stats.py``). It is the checker, never the
thing measured or shipped: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it. The product package
``sigpyproc3_amd`` never imports it and has no CPU fallback.

Modules
-------
``spp_oracle``  numpy restatement (every function cites the reference file:line).
``c_oracle``    ctypes loader for ``dedisp_omp.c`` (C/OpenMP restatement of
                ``kernels.dedisperse_par``), used for larger parity cases and as the
                CPU baseline.

Pinning status of each function is listed in the header of ``spp_oracle``.
"""
