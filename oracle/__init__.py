"""TEST INFRASTRUCTURE ONLY — the correctness oracle for the MI355X contraction backend.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this
package, and only as the checker or the timed CPU baseline.  The product path never imports
it and has no CPU fallback.

Contents
  loop.py            ctypes binding of ``loop_einsum.c`` — the brute-force restatement of
                     OMEinsum.loop_einsum! (src/loop_einsum.jl:21-45, src/Core.jl:122-268).
  default_backend.py numpy restatement of OMEinsum's DefaultBackend algorithm
                     (src/einsum.jl, src/unaryrules.jl, src/binaryrules.jl, src/utils.jl,
                     src/einsequence.jl, src/slicing.jl, src/autodiff.jl, src/bp.jl).

Conventions: arrays are numpy arrays whose axis p carries label ix[p] (the same convention
as Julia's ``size(x, p)``).  Inside the oracle every tensor is held in Fortran (column-major)
order so that Julia's free ``reshape`` stays free and the timed baseline has the reference's
cost structure.

How each function is pinned (beyond restating the reference text):
  * loop_einsum          closed forms of test/einsum.jl:103-212, the known answers
                         ``ein"ijk,ijk->"(s,s) == 6`` and Petersen == 0
                         (docs/src/applications.md:20-34), issue 136 (test/einsum.jl:306-309),
                         numpy.einsum on unique-label codes, brute-force sums on tiny inputs.
  * analyze_binary       fixture test/binaryrules.jl:5-16.
  * default einsum       against loop_einsum on the reference's code lists
                         (test/binaryrules.jl:18-81, test/einsum.jl:27-77,
                         test/roceinsum.jl:74-156) and the closed forms.
  * nested / sliced      nested == flat (test/einsequence.jl:28-48), sliced == unsliced
                         (test/slicing.jl:13-32), fullerene tropical value 66
                         (test/contractionorder.jl:31-70).
  * gradients            closed form ``einsum_grad(ij,jk) == c*transpose(b)``
                         (src/autodiff.jl:14-21), central finite differences, and
                         Petersen vertex-gradient sum == 0 (docs/src/applications.md:50-55).
"""
