"""Oracle: a CPU restatement of the reference FDTD s-parameter hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import anything under ``oracle/``, and
only as the checker (or the timed CPU baseline).  The product package
``oidt_amd`` never imports it and has no CPU fallback.

What it restates (paths relative to the reference root):

* ``oracle.setup``    -- ``FDTD/GPU/FDTD_2D_TM/s-parameter_extraction/fdtd_auto_setup.py``
  (and the TE / 3D twins), the CPML grading loops of ``fdtd_main.py`` and the
  source waveforms, with numpy-1.20 scalar promotion made explicit.
* ``oracle.geometry`` -- ``generate_fg_mask`` / ``apply_mask`` /
  ``average_materials`` (``fdtd_funcs.py``), given the roughness profile arrays
  (the reference's ``mode='load'`` path).
* ``oracle.fdtd``     -- every ``@cuda.jit`` kernel of the three bodies, in the
  reference launch order, as FP64 NumPy slicing.
* ``oracle.extract``  -- ``extract_sparams.py`` (2D and 3D) and the skrf-0.18.1
  Touchstone text layout.

How each part is pinned (the reference ships no tests or golden vectors):

* setup integers          -- grid sizes, nt and port columns derived from the
  reference's ``user_config.py`` defaults, recorded independently
  (``tests/golden/reference_setup_integers.json``).
* CPML / source / phasors -- closed-form properties (b in (0,1), kappa grading
  end points, Gaussian spectrum of the wave packet) and the CPML reflection
  test (an enlarged-domain run).
* geometry                -- hand-worked float-compare cases of ``apply_mask``
  and averaging on tiny grids; brute-force per-cell loops.
* curl / CPML / source / DFT kernels -- physics: zero source gives zero
  fields, linearity, discrete div(H)=0 in 3D, the slab-mode effective index
  of ``aux_funcs.find_neff`` recovered from the port DFT phase, CPML
  reflection below -40 dB against an enlarged domain, forward/backward mirror
  symmetry, passivity; plus a per-cell pure-Python loop transcription on tiny
  grids.
* extraction              -- synthetic phasors with a closed-form answer, and
  the assembly rules (upper triangle in 2D, lower triangle wins in 3D).

Parity unpinned (no reference output exists and the reference cannot run
here): ``pyspeckle`` profile generation and ``autocorrelation``; skrf's float
formatting beyond the example files' layout.
"""
