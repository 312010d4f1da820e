"""CPU oracle for the minkasi mapmaking inner loop -- TEST INFRASTRUCTURE ONLY.

This package restates, in FP64 on the CPU, the reference algorithms that the
MI355X build replaces.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
the timed CPU baseline.  Nothing under ``minkasi_amd/`` imports it: the
product path runs the HIP library and raises if that library is missing.

Modules:
  ``oracle.core``    -- C projections + CAR pixelization (liboracle.so),
                        WCS construction, map sizing.
  ``oracle.noise``   -- DCT-I, smoothing, NoiseSmoothedSVD init/apply.
  ``oracle.solve``   -- Tod.dot / TodVec.dot / make_rhs / make_hits / run_pcg.
  ``oracle.learned`` -- learned-preconditioner forward/backward (no reference
                        counterpart).

Pinning status of every function is stated in its module header.
"""
