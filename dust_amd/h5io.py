"""DUST's HDF5 result files on libdust_h5.so (include/dust_h5.h).

    File(path, mode)          new_hdf5_file / open_hdf5_file and the
                              write_hdf5 / read_hdf5 helpers of
                              src/io/mod_hdf5_io.f90
    result_name(base, it)     <basename>_res_<it, I4.4>.h5 (save_status,
                              src/io/mod_dust_io.f90:160-162)
    save_status(path, ...)    the groups and datasets of save_status
                              (src/io/mod_dust_io.f90:127-437) that the hot
                              path produces: time, Parameters, Components
                              (Vort and the optional element results),
                              PanelWake, RingWake, ParticleWake, References
    read_wake(path)           the datasets load_wake reads
                              (src/aeroel/mod_wake.f90:597-778)
    read_solution(path)       the component intensities of load_solution
                              (src/io/mod_dust_io.f90:444-543)

Arrays are written in the file's order, which is the reference's with the
dimensions reversed: a Fortran x(3,n) is a C-ordered (n, 3) array here, as
h5py shows the reference's files.  Logicals are the one-character strings
'T' / 'F' the HDF5 Fortran layer writes.  There is no fallback: without
libdust_h5.so every call raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
H5_PATH = os.path.join(_HERE, "libdust_h5.so")

F64, I32, STR = 0, 1, 2
MAX_RANK = 8
MAX_CHAR_LEN = 255  # character(len=max_char_len) of the reference's string attributes

_i64p = C.POINTER(C.c_int64)
SIGNATURES = [
    ("dust_h5_last_error", C.c_char_p, []),
    ("dust_h5_create", C.c_int, [C.c_char_p, _i64p]),
    ("dust_h5_open", C.c_int, [C.c_char_p, _i64p]),
    ("dust_h5_close", C.c_int, [C.c_int64]),
    ("dust_h5_write", C.c_int, [C.c_int64, C.c_char_p, C.c_int, C.c_int, _i64p, C.c_void_p]),
    ("dust_h5_write_string", C.c_int, [C.c_int64, C.c_char_p, C.c_char_p]),
    ("dust_h5_write_attr", C.c_int,
     [C.c_int64, C.c_char_p, C.c_int, C.c_int, _i64p, C.c_void_p]),
    ("dust_h5_write_attr_string", C.c_int, [C.c_int64, C.c_char_p, C.c_char_p, C.c_int64]),
    ("dust_h5_exists", C.c_int, [C.c_int64, C.c_char_p, C.POINTER(C.c_int)]),
    ("dust_h5_info", C.c_int,
     [C.c_int64, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), _i64p]),
    ("dust_h5_read", C.c_int, [C.c_int64, C.c_char_p, C.c_int, C.c_void_p, C.c_int64]),
    ("dust_h5_read_string", C.c_int, [C.c_int64, C.c_char_p, C.c_char_p, C.c_int64]),
]

_lib = None


class H5Error(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(H5_PATH):
            raise RuntimeError(f"{H5_PATH} is missing: build it with dust_amd.build.build_h5() "
                               "(needs the HDF5 C library)")
        L = C.CDLL(H5_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def _check(rc, where):
    if rc != 0:
        raise H5Error(f"{where}: {lib().dust_h5_last_error().decode()} (status {rc})")


def _dims(shape):
    return (C.c_int64 * max(1, len(shape)))(*shape)


def result_name(basename, it):
    """File name of the it-th output (write(sit,'(I4.4)') it)."""
    return f"{basename}_res_{int(it):04d}.h5"


class File:
    """An HDF5 file opened for writing ('w', truncates) or reading ('r')."""

    def __init__(self, path, mode="r"):
        h = C.c_int64()
        if mode == "w":
            _check(lib().dust_h5_create(os.fsencode(path), C.byref(h)), f"create {path}")
        elif mode == "r":
            _check(lib().dust_h5_open(os.fsencode(path), C.byref(h)), f"open {path}")
        else:
            raise ValueError("mode is 'r' or 'w'")
        self.h = h.value
        self.path = path

    def close(self):
        if self.h is not None:
            _check(lib().dust_h5_close(self.h), "close")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _typed(a):
        a = np.asarray(a)
        # asarray with order="C" keeps 0-d values 0-d (scalar dataspace)
        if a.dtype.kind == "f":
            return np.asarray(a, dtype=np.float64, order="C"), F64
        if a.dtype.kind in "iub":
            return np.asarray(a, dtype=np.int32, order="C"), I32
        raise TypeError(f"unsupported dtype {a.dtype}")

    def write(self, name, a):
        """write_hdf5 of a number or an array (float -> F64LE, int -> I32LE;
        0-d -> scalar dataspace)."""
        a, t = self._typed(a)
        _check(lib().dust_h5_write(self.h, name.encode(), t, a.ndim, _dims(a.shape),
                                   a.ctypes.data), f"write {name}")

    def write_string(self, name, s):
        _check(lib().dust_h5_write_string(self.h, name.encode(), str(s).encode()),
               f"write {name}")

    def attr(self, name, value, length=0):
        """write_hdf5_attr on the root group: logicals as 'T'/'F', strings
        space-padded to `length` (0 = their own length), numbers and arrays
        as F64 / I32."""
        if isinstance(value, (bool, np.bool_)):
            value = "T" if value else "F"
        if isinstance(value, str):
            _check(lib().dust_h5_write_attr_string(self.h, name.encode(), value.encode(),
                                                   int(length)), f"attribute {name}")
            return
        a, t = self._typed(value)
        _check(lib().dust_h5_write_attr(self.h, name.encode(), t, a.ndim, _dims(a.shape),
                                        a.ctypes.data), f"attribute {name}")

    def exists(self, name):
        e = C.c_int()
        _check(lib().dust_h5_exists(self.h, name.encode(), C.byref(e)), f"exists {name}")
        return bool(e.value)

    def info(self, name):
        t, r = C.c_int(), C.c_int()
        d = (C.c_int64 * MAX_RANK)()
        _check(lib().dust_h5_info(self.h, name.encode(), C.byref(t), C.byref(r), d),
               f"info {name}")
        return t.value, tuple(d[i] for i in range(r.value))

    def read(self, name):
        """read_hdf5(_al): the whole dataset (numpy float64 / int32 array of
        the file's shape, or a str)."""
        t, shape = self.info(name)
        if t == STR:
            buf = C.create_string_buffer(shape[0] + 1)
            _check(lib().dust_h5_read_string(self.h, name.encode(), buf, shape[0] + 1),
                   f"read {name}")
            return buf.value.decode()
        out = np.zeros(shape, dtype=np.float64 if t == F64 else np.int32)
        _check(lib().dust_h5_read(self.h, name.encode(), t, out.ctypes.data, out.size),
               f"read {name}")
        return out


def save_status(path, time, params=None, components=(), panel_wake=None, particles=None,
                ring_wake_len=0, references=()):
    """Writes one result file in save_status's layout.

    params: sim_param values written as root attributes (save_sim_param,
        src/general/mod_sim_param.f90:897-976) and, for u_inf, P_inf,
        rho_inf, mu_inf, a_inf, also into the Parameters group (:171-177).
    components: dicts with comp_id, name, ref_tag, ref_id, vort (ne,) and
        optionally pres (ne,), dforce (ne, 3), dmom (ne, 3), surf_vel (ne, 3),
        didou_dt (ne,) -- the Components/CompNNN/Solution datasets (:182-318);
        Hinges/n_hinges is 0 (no hinges on this path).
    panel_wake: dict points (len+1, n_points, 3), vels (same), start_points
        (n_stripes, 2) 1-based, vort (len, n_stripes) (:351-358).
    ring_wake_len: rin_wake_len of an empty ring wake (no actuator disks on
        this path): RingWake with zero disks (:360-389).
    particles: dict points (n, 3), vels (n, 3), vort (n, 3) = dir mag,
        turbvisc (n,), vortex_rad (n,), last_pan_idou (n_stripes,) (:391-413).
    references: dicts with tag, offset (3,), R (3, 3), vel (3,), rot_vel
        (3, 3), rel_pol_pos (3,), rel_rot_pos (:415-436).
    """
    params = dict(params or {})
    with File(path, "w") as f:
        for k, v in params.items():
            if isinstance(v, str):
                f.attr(k, v, MAX_CHAR_LEN if len(v) > 1 else 0)
            else:
                f.attr(k, v)
        f.write("time", np.float64(time))
        for k in ("u_inf", "P_inf", "rho_inf", "mu_inf", "a_inf"):
            if k in params:
                f.write("Parameters/" + k, np.asarray(params[k], dtype=np.float64))
        f.write("Components/NComponents", np.int32(len(components)))
        for c in components:
            g = "Components/Comp%03d/" % int(c["comp_id"])
            f.write_string(g + "CompName", c["name"])
            f.write_string(g + "RefTag", c.get("ref_tag", "0"))
            f.write(g + "RefId", np.int32(c.get("ref_id", 0)))
            f.write(g + "Solution/Vort", np.asarray(c["vort"], dtype=np.float64))
            for key, dset in (("pres", "Pres"), ("dforce", "dF"), ("dmom", "dMom"),
                              ("surf_vel", "surf_vel"), ("didou_dt", "didou_dt")):
                if c.get(key) is not None:
                    f.write(g + "Solution/" + dset, np.asarray(c[key], dtype=np.float64))
            f.write(g + "Hinges/n_hinges", np.int32(0))
        if panel_wake is not None:
            pw = panel_wake
            f.write("PanelWake/WakePoints", pw["points"])
            f.write("PanelWake/WakeVels", pw["vels"])
            f.write("PanelWake/StartPoints", np.asarray(pw["start_points"], dtype=np.int32))
            f.write("PanelWake/WakeVort", pw["vort"])
        L = int(ring_wake_len)
        f.write("RingWake/WakePoints", np.zeros((L, 0, 3)))
        f.write("RingWake/Conn_pe", np.zeros(0, dtype=np.int32))
        f.write("RingWake/WakeCenters", np.zeros((L, 0, 3)))
        f.write("RingWake/WakeVort", np.zeros((L, 0)))
        if particles is not None:
            p = particles
            f.write("ParticleWake/WakePoints", p["points"])
            f.write("ParticleWake/WakeVels", p["vels"])
            f.write("ParticleWake/turbvisc", p["turbvisc"])
            f.write("ParticleWake/VortexRad", p["vortex_rad"])
            f.write("ParticleWake/WakeVort", p["vort"])
            f.write("ParticleWake/LastPanIdou", p["last_pan_idou"])
        f.write("References/NReferences", np.int32(len(references)))
        for i, r in enumerate(references):
            g = "References/Ref%03d/" % i
            f.write_string(g + "Tag", r["tag"])
            f.write(g + "Offset", np.asarray(r["offset"], dtype=np.float64))
            f.write(g + "R", np.asarray(r["R"], dtype=np.float64))
            f.write(g + "Vel", np.asarray(r["vel"], dtype=np.float64))
            f.write(g + "RotVel", np.asarray(r["rot_vel"], dtype=np.float64))
            f.write(g + "RelativePolPos", np.asarray(r["rel_pol_pos"], dtype=np.float64))
            f.write(g + "RelativeRotPos", np.float64(r["rel_rot_pos"]))


def read_wake(path):
    """The datasets load_wake reads (mod_wake.f90:611-720): PanelWake
    points, velocities, start points, intensities; RingWake connectivity;
    ParticleWake points, velocities (None when the file has none, as
    check_dset_hdf5 allows), vorticity vectors, last intensities, core
    radii; plus the file's time."""
    with File(path, "r") as f:
        out = dict(time=float(f.read("time")))
        out["pan_points"] = f.read("PanelWake/WakePoints")
        out["pan_vels"] = f.read("PanelWake/WakeVels")
        out["start_points"] = f.read("PanelWake/StartPoints")
        out["pan_vort"] = f.read("PanelWake/WakeVort")
        out["conn_pe"] = f.read("RingWake/Conn_pe")
        out["rin_vort"] = f.read("RingWake/WakeVort")
        out["points"] = f.read("ParticleWake/WakePoints")
        out["vels"] = (f.read("ParticleWake/WakeVels") if f.exists("ParticleWake/WakeVels")
                       else None)
        out["vort"] = f.read("ParticleWake/WakeVort")
        out["last_pan_idou"] = f.read("ParticleWake/LastPanIdou")
        out["vortex_rad"] = f.read("ParticleWake/VortexRad")
    return out


def read_solution(path, comp_id):
    """Components/CompNNN/Solution/Vort of load_solution
    (mod_dust_io.f90:444-543)."""
    with File(path, "r") as f:
        return f.read("Components/Comp%03d/Solution/Vort" % int(comp_id))


def particles_from_file(vort):
    """mag and dir of load_wake (mod_wake.f90:742-748): mag = norm2(vort);
    dir = vort / mag when mag > 1e-13, else vort itself."""
    vort = np.asarray(vort, dtype=np.float64)
    mag = np.sqrt(np.sum(vort * vort, axis=1))
    big = mag > 1e-13
    d = vort.copy()
    d[big] = vort[big] / mag[big, None]
    return mag, d
