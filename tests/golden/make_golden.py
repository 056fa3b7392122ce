"""Writes the oracle's golden vectors (tests/golden/*.npz).

Cases: the reference's pyramid and torus tet meshes with LevelSet's default
initial values and velocity (src/LevelSet.cu:214-231), 3 steps each, and the
2D circle workload (17x17 jittered triangle grid, 10 steps). Each file holds,
per precision (FP32 and FP64 oracle instances), phi after every step and the
narrowband patch list after every step, the vertices whose advection update
hinges on a threshold decision near its rounding window (the oracle's
ambiguity flags), plus the inputs. Regenerate with
    python tests/golden/make_golden.py
after a deliberate change to the oracle."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def mesh_case(name, metis, steps=3):
    import levelset_amd.cases as cases
    xyz, tets = cases.read_tetgen(os.path.join(HERE, "meshes", name))
    tets = cases.reorient_tets(xyz, tets)
    phi0, vel = cases.levelset_default_x(xyz, tets.shape[0], 3)
    nparts = cases.metis_parts(xyz.shape[0], metis, 3)
    labels = cases.partition_rcb(xyz, nparts)
    return dict(dim=3, xyz=xyz, cells=tets, phi0=phi0.astype(np.float64), vel=vel, labels=labels,
                nparts=nparts, dt=1.0, bandwidth=16.0, niter=5, steps=steps)


def circle_case():
    import levelset_amd.cases as cases
    c = cases.config1()
    return dict(dim=2, xyz=c.xyz, cells=c.cells, phi0=c.phi0, vel=c.vel, labels=c.labels, nparts=c.nparts,
                dt=c.dt, bandwidth=c.bandwidth, niter=c.niter, steps=c.steps)


CASES = {"pyramid": lambda: mesh_case("pyramid", 16), "torus": lambda: mesh_case("torus", 32),
         "circle2d": circle_case}


def run_oracle(O, case, precision):
    o = O.Oracle(case["dim"], case["xyz"], case["cells"], case["labels"], case["nparts"], precision)
    o.set_phi(case["phi0"])
    o.set_velocity(case["vel"])
    prm = O.make_params(case["dt"], case["bandwidth"], case["niter"])
    phis, nbs, amb = [], [], []
    for _ in range(case["steps"]):
        st = o.step(prm)
        phis.append(o.phi())
        nbs.append(o.narrowband())
        amb.append(o.ambiguous())
        if st.nb_count == 0:
            break
    o.close()
    return np.stack(phis), nbs, np.stack(amb)


def main():
    from __graft_entry__ import load_package
    load_package()
    import oracle as O
    for name, make in CASES.items():
        case = make()
        out = {k: np.asarray(v) for k, v in case.items()}
        for prec in (32, 64):
            phis, nbs, amb = run_oracle(O, case, prec)
            out[f"phi_fp{prec}"] = phis
            out[f"amb_fp{prec}"] = amb
            out[f"nb_fp{prec}"] = np.concatenate(nbs).astype(np.int32)
            out[f"nb_len_fp{prec}"] = np.array([len(n) for n in nbs], dtype=np.int32)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"wrote {path}: {out['phi_fp64'].shape[0]} steps, {out['xyz'].shape[0]} vertices")


if __name__ == "__main__":
    main()
