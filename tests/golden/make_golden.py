"""Extract golden input vectors from the reference's own regression fixtures.

Reads (data only, no reference code is run) the HDF5 result files that DUST's
regression tests hold and writes small .npz fixtures next to this script:

  reg_test_b_particles.npz  -- ParticleWake of tests/reg_test_b (the only FMM
                               case: fmm=T, multipole_degree=2, 6 levels,
                               min_octree_part=5, box 6^3 at (-3,-3,-5),
                               vortex_rad 0.05, rankine_rad 0.005; from
                               tests/reg_test_b/dust_base.in)
  mir_static_p_particles.npz -- ParticleWake of tests/mir_test/static/p (the
                               direct-sum case with stretching and diffusion)
  reference_h5_layout.json  -- the dataset paths, shapes (as h5py reports
                               them) and element types of three result files
                               (reg_test_a static, reg_test_b, mir_test
                               static/p): the layout save_status writes

Run with an interpreter that has h5py (e.g. /opt/conda/bin/python3.9):
    /opt/conda/bin/python3.9 tests/golden/make_golden.py /root/reference
"""
import json
import os
import sys

import h5py
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def particles(path):
    with h5py.File(path, "r") as f:
        g = f["ParticleWake"]
        out = {"cen": np.asarray(g["WakePoints"][()], dtype=np.float64),
               "alpha": np.asarray(g["WakeVort"][()], dtype=np.float64)}
        if "VortexRad" in g:
            out["r_vortex"] = np.asarray(g["VortexRad"][()], dtype=np.float64)
        return out


def layout(path):
    """{dataset path: [shape, dtype]} of one result file, plus the root
    attributes' names."""
    out = {}
    with h5py.File(path, "r") as f:
        def visit(name, obj):
            if isinstance(obj, h5py.Dataset):
                out[name] = [list(obj.shape), obj.dtype.str]
        f.visititems(visit)
        attrs = sorted(f.attrs.keys())
    return {"datasets": out, "attrs": attrs}


def main(ref):
    b = particles(os.path.join(ref, "tests/reg_test_b/Output/ref_base_0-6-0_res_0011.h5"))
    np.savez_compressed(os.path.join(HERE, "reg_test_b_particles.npz"), **b,
                        box_length=6.0, nbox=np.array([1, 1, 1]),
                        origin=np.array([-3.0, -3.0, -5.0]), nlevels=6, min_part=5, degree=2,
                        vortex_rad=0.05, rankine_rad=0.005)
    m = particles(os.path.join(ref, "tests/mir_test/static/p/Output/ref_res_0011.h5"))
    np.savez_compressed(os.path.join(HERE, "mir_static_p_particles.npz"), **m, vortex_rad=0.1)
    files = {"reg_test_a_static": "tests/reg_test_a/Output/ref_static_base_0-6-0_res_0011.h5",
             "reg_test_b": "tests/reg_test_b/Output/ref_base_0-6-0_res_0011.h5",
             "mir_test_static_p": "tests/mir_test/static/p/Output/ref_res_0011.h5"}
    lay = {k: layout(os.path.join(ref, v)) for k, v in files.items()}
    with open(os.path.join(HERE, "reference_h5_layout.json"), "w") as fo:
        json.dump({"files": files, "layout": lay}, fo, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
