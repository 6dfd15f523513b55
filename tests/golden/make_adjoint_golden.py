"""Write adjoint_<name>.npz: the gradient of a gradient scenario's objective
computed by the reference package itself, stepping its torch float64 CPU
backend with autograd enabled (flaport/fdtd 0.3.6 has no other
differentiation path). Scenarios: g5 (CPML on all faces) and pa (periodic in
z, an anisotropic object with a full 3x3 inverse permittivity).

Run once, in the build container only:

    python tests/golden/make_adjoint_golden.py /root/reference [g5] [pa]

Contents: permittivity (the seeded design values, also regenerated from
adjoint_scenarios.py), J, grad = dJ/d permittivity, and the final E of the
run (for the forward check).
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import adjoint_scenarios as A  # noqa: E402


def main(ref_root, names):
    sys.path.insert(0, ref_root)
    import torch
    import fdtd

    fdtd.set_backend("torch.float64")
    torch.set_grad_enabled(True)  # the reference's torch backend disables it on import
    for name in names or ["g5", "pa"]:
        write(fdtd, torch, {"g5": A.G5, "pa": A.PA}[name])


def write(fdtd, torch, sc):
    eps0 = A.design_permittivity(sc)
    perm = torch.tensor(eps0, dtype=torch.float64, requires_grad=True)
    grid, design = A.build(fdtd, sc, perm)
    J = torch.zeros((), dtype=torch.float64)
    for _ in range(sc["steps"]):
        grid.step()
        # clone: E is updated in place by the next step
        J = J + (grid.E[sc["plane"]].clone() ** 2).sum()
    J.backward()
    out = os.path.join(HERE, f"adjoint_{sc['name']}.npz")
    np.savez_compressed(out, permittivity=eps0, J=np.array(J.item()), grad=perm.grad.numpy(),
                        E=grid.E.detach().numpy())
    print("wrote", out, "J =", J.item(), "|grad| =", float(perm.grad.norm()))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference", sys.argv[2:])
