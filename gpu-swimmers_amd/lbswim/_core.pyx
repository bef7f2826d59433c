# cython: language_level=3, boundscheck=False, wraparound=False
"""Cython bindings of the lbswim C ABI (include/lbswim.h).

Python classes with the reference package's names and behaviour:
  lb.Lattice, lb.LBParams, lb.LatticeAddressing, lb.SharedLatticeArray
      (lbswim/lb.pyx:11-245)
  shared.Array, shared.PrngArray                    (lbswim/shared.pyx:1-66)
  swimmers.CommonParams, swimmers.Array             (lbswim/swimmers.pyx:1-155)
  tracers.Array                                     (lbswim/tracers.pyx:1-57)
numpy arrays alias the library's FP64 host mirrors, as in the reference; H2D()
and D2H() move whole fields.  Errors raise instead of exiting the process.
"""
cimport cython
from libc.stdint cimport uint64_t
cimport numpy as cnp
import numpy as np

cnp.import_array()

cdef extern from "lbswim.h" nogil:
    int LBSWIM_OK, LBSWIM_EINVAL, LBSWIM_ENOMEM, LBSWIM_EHIP, LBSWIM_ENODEV
    int LBSWIM_STORE_F64, LBSWIM_STORE_F32DEV
    int LBSWIM_FORCE_STORED, LBSWIM_FORCE_AUTO, LBSWIM_FORCE_VIRTUAL, LBSWIM_FORCE_SCATTER
    int LBSWIM_RHO, LBSWIM_U, LBSWIM_FORCE, LBSWIM_FOLD, LBSWIM_FNEW
    int LBSWIM_PART_R, LBSWIM_PART_V, LBSWIM_PART_N, LBSWIM_PART_S, LBSWIM_PART_PRNG
    int LBSWIM_K_COUNT

    ctypedef struct lbswim_params:
        double cs2
        double w[15]
        double xi[15][3]
        int ci[15][3]
        double Q[15][3][3]
        int complement[15]
        double norms[15]
        double mm[15][15]
        double mmi[15][15]
        double tau_s
        double tau_b

    ctypedef struct lbswim_addressing:
        int size[3]
        int strides[3]
        int n

    ctypedef struct lbswim_common_params:
        double P
        double l
        double alpha
        double mobility
        unsigned char translational_advection_off
        unsigned char rotational_advection_off
        unsigned long long seed

    ctypedef struct lbswim_lattice:
        pass
    ctypedef struct lbswim_swimmers:
        pass
    ctypedef struct lbswim_tracers:
        pass

    const char* lbswim_last_error()
    int lbswim_synchronize()
    ctypedef struct lbswim_mark:
        pass
    int lbswim_mark_create(lbswim_mark** out)
    int lbswim_mark_wait(lbswim_mark* m)
    int lbswim_mark_reached(lbswim_mark* m, int* reached)
    int lbswim_mark_elapsed(lbswim_mark* a, lbswim_mark* b, double* ms)
    void lbswim_mark_destroy(lbswim_mark* m)
    int lbswim_lattice_create(int nx, int ny, int nz, double tau_s, double tau_b, int storage,
                              lbswim_lattice** out)
    void lbswim_lattice_destroy(lbswim_lattice* lat)
    int lbswim_lattice_step(lbswim_lattice* lat)
    int lbswim_lattice_calc_hydro(lbswim_lattice* lat)
    int lbswim_lattice_init_from_hydro(lbswim_lattice* lat)
    int lbswim_lattice_zero_force(lbswim_lattice* lat)
    int lbswim_lattice_set_force_mode(lbswim_lattice* lat, int mode)
    int lbswim_lattice_get_force_mode(lbswim_lattice* lat, int* mode)
    int lbswim_lattice_sync_force(lbswim_lattice* lat)
    int lbswim_lattice_force_is_virtual(lbswim_lattice* lat, int* is_virtual)
    int lbswim_lattice_sync_u(lbswim_lattice* lat)
    int lbswim_lattice_u_is_virtual(lbswim_lattice* lat, int* is_virtual)
    int lbswim_lattice_set_fusion(lbswim_lattice* lat, int enable, int rows, int strips)
    int lbswim_lattice_fusion(lbswim_lattice* lat, int* active, int* rows, int* strips)
    int lbswim_lattice_params(lbswim_lattice* lat, const lbswim_params** host)
    int lbswim_lattice_addressing(lbswim_lattice* lat, const lbswim_addressing** host)
    int lbswim_lattice_get_time_step(lbswim_lattice* lat, int* out)
    int lbswim_lattice_set_time_step(lbswim_lattice* lat, int ts)
    int lbswim_lattice_storage(lbswim_lattice* lat, int* out)
    int lbswim_lattice_field(lbswim_lattice* lat, int field, double** host, void** device, size_t* size)
    int lbswim_lattice_h2d(lbswim_lattice* lat, int field)
    int lbswim_lattice_d2h(lbswim_lattice* lat, int field)
    int lbswim_swimmers_create(int n, const lbswim_common_params* p, lbswim_swimmers** out)
    void lbswim_swimmers_destroy(lbswim_swimmers* sw)
    int lbswim_swimmers_common(lbswim_swimmers* sw, lbswim_common_params** host)
    int lbswim_swimmers_common_h2d(lbswim_swimmers* sw)
    int lbswim_swimmers_add_forces(lbswim_swimmers* sw, lbswim_lattice* lat)
    int lbswim_swimmers_move(lbswim_swimmers* sw, lbswim_lattice* lat)
    int lbswim_swimmers_field(lbswim_swimmers* sw, int field, void** host, void** device, size_t* count,
                              size_t* elem_size)
    int lbswim_swimmers_h2d(lbswim_swimmers* sw, int field)
    int lbswim_swimmers_d2h(lbswim_swimmers* sw, int field)
    int lbswim_swimmers_tumbles(lbswim_swimmers* sw, unsigned long long* out)
    int lbswim_tracers_create(int n, lbswim_tracers** out)
    void lbswim_tracers_destroy(lbswim_tracers* tr)
    int lbswim_tracers_move(lbswim_tracers* tr, lbswim_lattice* lat)
    int lbswim_tracers_field(lbswim_tracers* tr, int field, void** host, void** device, size_t* count)
    int lbswim_tracers_h2d(lbswim_tracers* tr, int field)
    int lbswim_tracers_d2h(lbswim_tracers* tr, int field)
    int lbswim_system_autotune(lbswim_lattice* lat, lbswim_swimmers** sw, int nsw, lbswim_tracers** tr, int ntr,
                               int steps_per_trial, int* rows, int* strips)
    int lbswim_system_step(lbswim_lattice* lat, lbswim_swimmers** sw, int nsw, lbswim_tracers** tr, int ntr,
                           int nsteps)
    ctypedef struct lbswim_ensemble:
        pass
    int lbswim_ensemble_create(int storage, lbswim_ensemble** out)
    void lbswim_ensemble_destroy(lbswim_ensemble* e)
    int lbswim_ensemble_add(lbswim_ensemble* e, lbswim_lattice* lat, lbswim_swimmers** sw, int nsw,
                            lbswim_tracers** tr, int ntr, int* member)
    int lbswim_ensemble_remove(lbswim_ensemble* e, int member)
    int lbswim_ensemble_size(lbswim_ensemble* e, int* members)
    int lbswim_ensemble_step(lbswim_ensemble* e, int nsteps)
    int lbswim_ensemble_set_graph(lbswim_ensemble* e, int on)
    int lbswim_ensemble_set_order(lbswim_ensemble* e, int order)
    int lbswim_ensemble_set_force_mode(lbswim_ensemble* e, int mode)
    int lbswim_ensemble_set_work(lbswim_ensemble* e, int blocks_per_item)
    int lbswim_ensemble_tuning(lbswim_ensemble* e, int* order, int* blocks_per_item)
    int lbswim_ensemble_set_timing(lbswim_ensemble* e, unsigned mask)
    int lbswim_ensemble_kernel_time(lbswim_ensemble* e, int kernel, double* total_ms, long* count)
    int lbswim_ensemble_sites(lbswim_ensemble* e, size_t* virtual_sites, size_t* stored_sites)
    int lbswim_ensemble_autotune(lbswim_ensemble* e, int steps_per_trial, int* choice)
    int lbswim_ensemble_set_fusion(lbswim_ensemble* e, int rows, int strip)
    int lbswim_ensemble_fusion(lbswim_ensemble* e, int* rows, int* strip)
    int lbswim_ensemble_set_lanes(lbswim_ensemble* e, int lanes)
    int lbswim_ensemble_lanes(lbswim_ensemble* e, int* lanes)
    int lbswim_ensemble_fused_sites(lbswim_ensemble* e, size_t* fused_sites)
    int lbswim_lattice_totals(lbswim_lattice* lat, double* out)
    int lbswim_lattice_set_timing(lbswim_lattice* lat, int on)
    int lbswim_lattice_set_timing_mask(lbswim_lattice* lat, unsigned mask)
    int lbswim_lattice_kernel_time(lbswim_lattice* lat, int kernel, double* total_ms, long* count)
    int lbswim_device_memory(size_t* free_bytes, size_t* total_bytes)
    int lbswim_diag_copy(size_t count, int reps, double* ms)
    int lbswim_diag_peskin(const double* x, double* out, int count)
    int lbswim_diag_interp(int nx, int ny, int nz, const double* u, const double* r, double* v, int count)


class LbswimError(RuntimeError):
    """A failure reported by the native library."""


cdef int check(int rc) except -1:
    if rc == LBSWIM_OK:
        return 0
    msg = lbswim_last_error().decode("utf-8", "replace")
    if rc == LBSWIM_EINVAL:
        raise ValueError(msg)
    if rc == LBSWIM_ENOMEM:
        raise MemoryError(msg)
    raise LbswimError(msg)


STORAGE = {"f64": LBSWIM_STORE_F64, "f32": LBSWIM_STORE_F32DEV}
KERNELS = ("collide_stream", "calc_hydro", "zero_force", "add_forces", "swimmers_move", "tracers_move",
           "init_from_hydro", "step_hydro", "stored_chain")


cdef object wrap_view(object owner, void* data, int ndim, cnp.npy_intp* shape, int typenum):
    """numpy view of library-owned memory that keeps `owner` alive."""
    arr = cnp.PyArray_SimpleNewFromData(ndim, shape, typenum, data)
    cnp.set_array_base(arr, owner)
    return arr


def synchronize():
    """Wait until all enqueued device work has finished (other Python
    threads run meanwhile)."""
    cdef int rc
    with nogil:
        rc = lbswim_synchronize()
    check(rc)


cdef class Mark:
    """A point in the library's stream: all device work enqueued before the
    mark was made.  wait() blocks until the device has reached it (other
    Python threads run meanwhile); reached tests it; elapsed_ms times the
    device between two marks."""
    cdef lbswim_mark* impl

    def __cinit__(self):
        self.impl = NULL
        check(lbswim_mark_create(&self.impl))

    def __dealloc__(self):
        if self.impl is not NULL:
            lbswim_mark_destroy(self.impl)
            self.impl = NULL

    def wait(self):
        cdef int rc
        with nogil:
            rc = lbswim_mark_wait(self.impl)
        check(rc)

    @property
    def reached(self):
        cdef int r = 0
        check(lbswim_mark_reached(self.impl, &r))
        return bool(r)

    def elapsed_ms(self, Mark later):
        """Device time from this mark to `later` (both reached)."""
        cdef double ms = 0
        check(lbswim_mark_elapsed(self.impl, later.impl, &ms))
        return ms


FORCE_MODES = ("stored", "auto", "virtual", "scatter")


cdef int force_mode_id(mode) except -1:
    ids = {"stored": LBSWIM_FORCE_STORED, "auto": LBSWIM_FORCE_AUTO, "virtual": LBSWIM_FORCE_VIRTUAL,
           "scatter": LBSWIM_FORCE_SCATTER}
    if mode not in ids:
        raise ValueError("force mode must be one of %s" % (FORCE_MODES,))
    return ids[mode]


# ---------------------------------------------------------------------------
# shared.Array / shared.PrngArray (lbswim/shared.pyx:1-66)
# ---------------------------------------------------------------------------

cdef class Array:
    """A host/device field pair: .data is the numpy view of the host mirror."""
    cdef object owner
    cdef object view
    cdef int field

    @property
    def data(self):
        return self.view

    def H2D(self):
        self.owner._field_h2d(self.field)

    def D2H(self):
        self.owner._field_d2h(self.field)

    def __getattr__(self, attr):
        return getattr(self.view, attr)

    def __getitem__(self, idx):
        return self.view[idx]

    def __setitem__(self, idx, val):
        self.view[idx] = val

    def __len__(self):
        return len(self.view)

    def __array__(self, dtype=None, copy=None):
        return np.asarray(self.view, dtype=dtype)


cdef Array make_array(object owner, object view, int field, cls=Array):
    cdef Array a = cls.__new__(cls)
    a.owner = owner
    a.view = view
    a.field = field
    return a


class PrngArray(Array):
    """Per-swimmer random-stream state (a uint64 draw counter per swimmer)."""


# ---------------------------------------------------------------------------
# lb module (lbswim/lb.pyx)
# ---------------------------------------------------------------------------

cdef class LatticeAddressing:
    cdef const lbswim_addressing* impl
    cdef object lat

    @property
    def size(self):
        return (self.impl.size[0], self.impl.size[1], self.impl.size[2])

    @property
    def strides(self):
        return (self.impl.strides[0], self.impl.strides[1], self.impl.strides[2])

    @property
    def n(self):
        return self.impl.n


cdef class LBParams:
    """The D3Q15 model parameters (lbswim/lb.pyx:57-137). Treat as read only."""
    cdef const lbswim_params* impl
    cdef object lat

    cdef _arr(self, const void* p, int ndim, cnp.npy_intp* shape, int typenum):
        arr = wrap_view(self, <void*>p, ndim, shape, typenum)
        arr.flags.writeable = False
        return arr

    @property
    def cs2(self):
        return self.impl.cs2

    @property
    def w(self):
        cdef cnp.npy_intp shape[1]
        shape[0] = 15
        return self._arr(self.impl.w, 1, shape, cnp.NPY_DOUBLE)

    @property
    def xi(self):
        cdef cnp.npy_intp shape[2]
        shape[0] = 15
        shape[1] = 3
        return self._arr(self.impl.xi, 2, shape, cnp.NPY_DOUBLE)

    @property
    def ci(self):
        cdef cnp.npy_intp shape[2]
        shape[0] = 15
        shape[1] = 3
        return self._arr(self.impl.ci, 2, shape, cnp.NPY_INT)

    @property
    def Q(self):
        cdef cnp.npy_intp shape[3]
        shape[0] = 15
        shape[1] = 3
        shape[2] = 3
        return self._arr(self.impl.Q, 3, shape, cnp.NPY_DOUBLE)

    @property
    def complement(self):
        cdef cnp.npy_intp shape[1]
        shape[0] = 15
        return self._arr(self.impl.complement, 1, shape, cnp.NPY_INT)

    @property
    def norms(self):
        cdef cnp.npy_intp shape[1]
        shape[0] = 15
        return self._arr(self.impl.norms, 1, shape, cnp.NPY_DOUBLE)

    @property
    def mm(self):
        cdef cnp.npy_intp shape[2]
        shape[0] = 15
        shape[1] = 15
        return self._arr(self.impl.mm, 2, shape, cnp.NPY_DOUBLE)

    @property
    def mmi(self):
        cdef cnp.npy_intp shape[2]
        shape[0] = 15
        shape[1] = 15
        return self._arr(self.impl.mmi, 2, shape, cnp.NPY_DOUBLE)

    @property
    def tau_s(self):
        return self.impl.tau_s

    @property
    def tau_b(self):
        return self.impl.tau_b

    @property
    def eta(self):
        """Viscosity of the fluid."""
        return self.tau_s * self.cs2


cdef class SharedLatticeArray:
    """A lattice field viewed with shape (k, nx, ny, nz) (lbswim/lb.pyx:21-46)."""
    cdef Array impl
    cdef object view

    def __init__(self, Lattice lat, Array sa):
        self.impl = sa
        n = lat.addr.n
        self.view = sa.view.reshape((sa.view.size // n,) + lat.addr.size)

    @property
    def data(self):
        return self.view

    def H2D(self):
        self.impl.H2D()

    def D2H(self):
        self.impl.D2H()

    def __getattr__(self, attr):
        return getattr(self.view, attr)

    def __getitem__(self, idx):
        return self.view[idx]

    def __setitem__(self, idx, val):
        self.view[idx] = val

    def __array__(self, dtype=None, copy=None):
        return np.asarray(self.view, dtype=dtype)


_FIELDS = {"rho": LBSWIM_RHO, "u": LBSWIM_U, "force": LBSWIM_FORCE, "fOld": LBSWIM_FOLD, "fNew": LBSWIM_FNEW}


cdef class Lattice:
    """Lattice(nx, ny, nz, tau_s, tau_b, storage="f64") (lbswim/lb.pyx:139-245).

    storage "f64" keeps the reference's FP64 arithmetic; "f32" stores the
    distributions as FP32 deviations from the rest state (the host views stay
    FP64 in both forms).
    """
    cdef lbswim_lattice* impl
    cdef dict _views
    cdef LBParams _params
    cdef LatticeAddressing _addr
    cdef str _storage

    def __cinit__(self, int nx, int ny, int nz, double tau_s, double tau_b, storage="f64"):
        self.impl = NULL
        if storage not in STORAGE:
            raise ValueError("storage must be 'f64' or 'f32', not %r" % (storage,))
        check(lbswim_lattice_create(nx, ny, nz, tau_s, tau_b, STORAGE[storage], &self.impl))
        self._views = {}
        self._storage = storage

    def __dealloc__(self):
        cdef lbswim_lattice* p = self.impl
        self.impl = NULL
        if p is not NULL:
            with nogil:
                lbswim_lattice_destroy(p)

    cdef lbswim_lattice* handle(self) except NULL:
        if self.impl is NULL:
            raise ValueError("lattice is not initialised")
        return self.impl

    def Step(self):
        check(lbswim_lattice_step(self.handle()))

    def CalcHydro(self):
        check(lbswim_lattice_calc_hydro(self.handle()))

    def InitFromHydro(self):
        check(lbswim_lattice_init_from_hydro(self.handle()))

    def ZeroForce(self):
        check(lbswim_lattice_zero_force(self.handle()))

    def set_force_mode(self, mode):
        """"auto" (default): the coupled step keeps the swimmer force virtual
        where swimmers are sparse (the force array is written on demand);
        "stored": always write it; "virtual": never write it in the step;
        "scatter" (FP32 storage): the swimmer moves add the next step's force
        into a private buffer with atomics, which the fused step reads (no
        spreading lists; sums rounded in a varying order, so runs agree to
        FP32 rounding rather than bit for bit)."""
        check(lbswim_lattice_set_force_mode(self.handle(), force_mode_id(mode)))

    @property
    def force_mode(self):
        """The force mode in use (System.Autotune may change it)."""
        cdef int m = 0
        check(lbswim_lattice_get_force_mode(self.handle(), &m))
        return FORCE_MODES[m]

    def sync_force(self):
        """Write the force array if the last coupled step left it virtual."""
        check(lbswim_lattice_sync_force(self.handle()))

    def sync_u(self):
        """Write u in the reference layout if fused steps left it only in the
        particle moves' component-interleaved copy."""
        check(lbswim_lattice_sync_u(self.handle()))

    @property
    def u_is_virtual(self):
        """True while u is not yet written in the reference layout."""
        cdef int v = 0
        check(lbswim_lattice_u_is_virtual(self.handle(), &v))
        return bool(v)

    def set_fusion(self, enable=True, int rows=0, int strips=0):
        """Fuse Step and CalcHydro of the coupled step into one launch.
        enable: True (the default) where the tiles are full (nz % 64 == 0,
        rows dividing ny), "always" on every lattice with ny >= rows, False
        never.  rows (2, 4, 8) and strips 0 choose the tile shape."""
        mode = 2 if enable == "always" else (1 if enable else 0)
        check(lbswim_lattice_set_fusion(self.handle(), mode, rows, strips))

    @property
    def fusion(self):
        """(active, rows, strips) of the coupled step's fused Step + CalcHydro."""
        cdef int a = 0, r = 0, s = 0
        check(lbswim_lattice_fusion(self.handle(), &a, &r, &s))
        return bool(a), r, s

    @property
    def force_is_virtual(self):
        """True while the last coupled step's force is not yet written to the force array."""
        cdef int v = 0
        check(lbswim_lattice_force_is_virtual(self.handle(), &v))
        return bool(v)

    @property
    def time_step(self):
        cdef int ts = 0
        check(lbswim_lattice_get_time_step(self.handle(), &ts))
        return ts

    @time_step.setter
    def time_step(self, int ts):
        check(lbswim_lattice_set_time_step(self.handle(), ts))

    @property
    def storage(self):
        return self._storage

    @property
    def addr(self):
        cdef const lbswim_addressing* a = NULL
        if self._addr is None:
            check(lbswim_lattice_addressing(self.handle(), &a))
            self._addr = LatticeAddressing.__new__(LatticeAddressing)
            self._addr.impl = a
            self._addr.lat = self
        return self._addr

    @property
    def params(self):
        cdef const lbswim_params* p = NULL
        if self._params is None:
            check(lbswim_lattice_params(self.handle(), &p))
            self._params = LBParams.__new__(LBParams)
            self._params.impl = p
            self._params.lat = self
        return self._params

    cdef object _field(self, str name):
        cdef double* host = NULL
        cdef size_t size = 0
        cdef cnp.npy_intp shape[1]
        cdef int fid = _FIELDS[name]
        if name not in self._views:
            check(lbswim_lattice_field(self.handle(), fid, &host, NULL, &size))
            shape[0] = <cnp.npy_intp>size
            view = wrap_view(self, host, 1, shape, cnp.NPY_DOUBLE)
            self._views[name] = SharedLatticeArray(self, make_array(self, view, fid))
        return self._views[name]

    def _field_h2d(self, int field):
        cdef lbswim_lattice* h = self.handle()
        cdef int rc
        with nogil:
            rc = lbswim_lattice_h2d(h, field)
        check(rc)

    def _field_d2h(self, int field):
        cdef lbswim_lattice* h = self.handle()
        cdef int rc
        with nogil:
            rc = lbswim_lattice_d2h(h, field)
        check(rc)

    @property
    def rho(self):
        return self._field("rho")

    @property
    def u(self):
        return self._field("u")

    @property
    def force(self):
        return self._field("force")

    @property
    def fOld(self):
        return self._field("fOld")

    @property
    def fNew(self):
        return self._field("fNew")

    def totals(self):
        """FP64 total mass and momentum of the current distributions (array of 4)."""
        cdef double out[4]
        check(lbswim_lattice_totals(self.handle(), out))
        return np.array([out[0], out[1], out[2], out[3]])

    def set_timing(self, on):
        """True/False: record every kernel or none; or an iterable of kernel
        names (from KERNELS) to record only those."""
        if on is True or on is False or on in (0, 1):
            check(lbswim_lattice_set_timing(self.handle(), 1 if on else 0))
            return
        mask = 0
        for name in on:
            mask |= 1 << KERNELS.index(name)
        check(lbswim_lattice_set_timing_mask(self.handle(), mask))

    def kernel_times(self):
        """{kernel: (total_ms, launches)} recorded since set_timing(True)."""
        cdef double ms = 0
        cdef long cnt = 0
        out = {}
        for k, name in enumerate(KERNELS):
            check(lbswim_lattice_kernel_time(self.handle(), k, &ms, &cnt))
            out[name] = (ms, cnt)
        return out

    def __reduce__(self):
        """Pickle as (class, ctor args, (time_step, fOld)) like lbswim/lb.pyx:221-235."""
        size = self.addr.size
        p = self.params
        self.fOld.D2H()
        data = (self.time_step, np.array(self.fOld.data, copy=True))
        return (Lattice, (size[0], size[1], size[2], p.tau_s, p.tau_b, self._storage), data)

    def __setstate__(self, data):
        ts, f = data
        self.time_step = ts
        self.fOld.data[:] = f[:]
        self.fOld.H2D()


# ---------------------------------------------------------------------------
# swimmers module (lbswim/swimmers.pyx)
# ---------------------------------------------------------------------------

cdef class CommonParams:
    """Swimmer parameters shared by one array (lbswim/swimmers.pyx:5-92).

    A free-standing instance owns its values; the one returned by
    swimmers.Array.common edits the array's parameters in place and pushes
    each change to the device, as the reference's setters do.
    """
    cdef lbswim_common_params own
    cdef lbswim_common_params* impl
    cdef lbswim_swimmers* sw
    cdef object sw_ref

    def __cinit__(self):
        self.own.P = 0
        self.own.l = 0
        self.own.alpha = 0
        self.own.mobility = 0
        self.own.translational_advection_off = 0
        self.own.rotational_advection_off = 0
        self.own.seed = 0
        self.impl = &self.own
        self.sw = NULL
        self.sw_ref = None

    cdef void adopt(self, lbswim_swimmers* sw, object owner):
        cdef lbswim_common_params* p = NULL
        lbswim_swimmers_common(sw, &p)
        self.impl = p
        self.sw = sw
        self.sw_ref = owner

    cdef void push(self):
        if self.sw is not NULL:
            lbswim_swimmers_common_h2d(self.sw)

    @property
    def P(self):
        return self.impl.P

    @P.setter
    def P(self, double v):
        self.impl.P = v
        self.push()

    @property
    def l(self):
        return self.impl.l

    @l.setter
    def l(self, double v):
        self.impl.l = v
        self.push()

    @property
    def alpha(self):
        return self.impl.alpha

    @alpha.setter
    def alpha(self, double v):
        self.impl.alpha = v
        self.push()

    @property
    def mobility(self):
        return self.impl.mobility

    @mobility.setter
    def mobility(self, double v):
        self.impl.mobility = v
        self.push()

    @property
    def translational_advection_off(self):
        return bool(self.impl.translational_advection_off)

    @translational_advection_off.setter
    def translational_advection_off(self, bint v):
        self.impl.translational_advection_off = v
        self.push()

    @property
    def rotational_advection_off(self):
        return bool(self.impl.rotational_advection_off)

    @rotational_advection_off.setter
    def rotational_advection_off(self, bint v):
        self.impl.rotational_advection_off = v
        self.push()

    @property
    def seed(self):
        return self.impl.seed

    @seed.setter
    def seed(self, unsigned long long v):
        self.impl.seed = v
        self.push()

    def as_dict(self):
        return {"P": self.P, "l": self.l, "alpha": self.alpha, "mobility": self.mobility,
                "translational_advection_off": self.translational_advection_off,
                "rotational_advection_off": self.rotational_advection_off, "seed": self.seed}

    def __reduce__(self):
        return (CommonParams, (), self.as_dict())

    def __setstate__(self, params):
        self.impl.P = params["P"]
        self.impl.l = params["l"]
        self.impl.alpha = params["alpha"]
        self.impl.mobility = params["mobility"]
        self.impl.translational_advection_off = params["translational_advection_off"]
        self.impl.rotational_advection_off = params["rotational_advection_off"]
        self.impl.seed = params["seed"]
        self.push()


cdef class SwimmerArray:
    """swimmers.Array(n, CommonParams) (lbswim/swimmers.pyx:94-155)."""
    cdef lbswim_swimmers* impl
    cdef int _num
    cdef Array _r, _v, _n, _prng
    cdef CommonParams _common

    def __cinit__(self, int n, CommonParams cp):
        self.impl = NULL
        if cp is None:
            raise TypeError("CommonParams required")
        check(lbswim_swimmers_create(n, cp.impl, &self.impl))
        self._num = n
        self._r = self._make(LBSWIM_PART_R, cnp.NPY_DOUBLE, Array)
        self._v = self._make(LBSWIM_PART_V, cnp.NPY_DOUBLE, Array)
        self._n = self._make(LBSWIM_PART_N, cnp.NPY_DOUBLE, Array)
        self._prng = self._make(LBSWIM_PART_PRNG, cnp.NPY_UINT64, PrngArray)
        self._common = CommonParams()
        self._common.adopt(self.impl, self)

    def __dealloc__(self):
        cdef lbswim_swimmers* p = self.impl
        self.impl = NULL
        if p is not NULL:
            with nogil:
                lbswim_swimmers_destroy(p)

    cdef Array _make(self, int field, int typenum, cls):
        cdef void* host = NULL
        cdef size_t count = 0, elem = 0
        cdef cnp.npy_intp shape[2]
        check(lbswim_swimmers_field(self.impl, field, &host, NULL, &count, &elem))
        if field == LBSWIM_PART_PRNG:
            shape[0] = self._num
            view = wrap_view(self, host, 1, shape, typenum)
        else:
            shape[0] = 3
            shape[1] = self._num
            view = wrap_view(self, host, 2, shape, typenum)
        return make_array(self, view, field, cls)

    def _field_h2d(self, int field):
        cdef lbswim_swimmers* h = self.handle()
        cdef int rc
        with nogil:
            rc = lbswim_swimmers_h2d(h, field)
        check(rc)

    def _field_d2h(self, int field):
        cdef lbswim_swimmers* h = self.handle()
        cdef int rc
        with nogil:
            rc = lbswim_swimmers_d2h(h, field)
        check(rc)

    cdef lbswim_swimmers* handle(self) except NULL:
        if self.impl is NULL:
            raise ValueError("swimmer array is not initialised")
        return self.impl

    cpdef AddForces(self, Lattice lat):
        check(lbswim_swimmers_add_forces(self.handle(), lat.handle()))

    cpdef Move(self, Lattice lat):
        check(lbswim_swimmers_move(self.handle(), lat.handle()))

    def H2D(self):
        self._r.H2D()
        self._v.H2D()
        self._n.H2D()
        self._prng.H2D()

    def D2H(self):
        self._r.D2H()
        self._v.D2H()
        self._n.D2H()
        self._prng.D2H()

    @property
    def num(self):
        return self._num

    @property
    def r(self):
        return self._r

    @property
    def v(self):
        return self._v

    @property
    def n(self):
        return self._n

    @property
    def prng(self):
        return self._prng

    @property
    def common(self):
        return self._common

    def tumbles(self):
        """Tumbles since creation (device counter)."""
        cdef unsigned long long t = 0
        check(lbswim_swimmers_tumbles(self.handle(), &t))
        return t

    def __reduce__(self):
        # (r, v, n) as lbswim/swimmers.pyx:143-148, plus the random-stream counters
        self.D2H()
        cp = CommonParams()
        cp.__setstate__(self._common.as_dict())
        data = (np.array(self._r.data), np.array(self._v.data), np.array(self._n.data),
                np.array(self._prng.data))
        return (SwimmerArray, (self._num, cp), data)

    def __setstate__(self, data):
        r, v, n = data[:3]
        self._r.data[:] = r[:]
        self._v.data[:] = v[:]
        self._n.data[:] = n[:]
        if len(data) > 3:
            self._prng.data[:] = data[3][:]
        self.H2D()


# ---------------------------------------------------------------------------
# tracers module (lbswim/tracers.pyx)
# ---------------------------------------------------------------------------

cdef class TracerArray:
    """tracers.Array(n) (lbswim/tracers.pyx:5-57)."""
    cdef lbswim_tracers* impl
    cdef int _num
    cdef Array _r, _s, _v

    def __cinit__(self, int n):
        self.impl = NULL
        check(lbswim_tracers_create(n, &self.impl))
        self._num = n
        self._r = self._make(LBSWIM_PART_R)
        self._s = self._make(LBSWIM_PART_S)
        self._v = self._make(LBSWIM_PART_V)

    def __dealloc__(self):
        cdef lbswim_tracers* p = self.impl
        self.impl = NULL
        if p is not NULL:
            with nogil:
                lbswim_tracers_destroy(p)

    cdef Array _make(self, int field):
        cdef void* host = NULL
        cdef size_t count = 0
        cdef cnp.npy_intp shape[2]
        check(lbswim_tracers_field(self.impl, field, &host, NULL, &count))
        shape[0] = 3
        shape[1] = self._num
        view = wrap_view(self, host, 2, shape, cnp.NPY_DOUBLE)
        return make_array(self, view, field)

    def _field_h2d(self, int field):
        cdef lbswim_tracers* h = self.handle()
        cdef int rc
        with nogil:
            rc = lbswim_tracers_h2d(h, field)
        check(rc)

    def _field_d2h(self, int field):
        cdef lbswim_tracers* h = self.handle()
        cdef int rc
        with nogil:
            rc = lbswim_tracers_d2h(h, field)
        check(rc)

    cdef lbswim_tracers* handle(self) except NULL:
        if self.impl is NULL:
            raise ValueError("tracer array is not initialised")
        return self.impl

    cpdef AddForces(self, Lattice lat):
        pass

    cpdef Move(self, Lattice lat):
        check(lbswim_tracers_move(self.handle(), lat.handle()))

    def H2D(self):
        self._r.H2D()
        self._s.H2D()
        self._v.H2D()

    def D2H(self):
        self._r.D2H()
        self._s.D2H()
        self._v.D2H()

    @property
    def num(self):
        return self._num

    @property
    def r(self):
        return self._r

    @property
    def s(self):
        return self._s

    @property
    def v(self):
        return self._v

    def __reduce__(self):
        self.D2H()
        data = (np.array(self._r.data), np.array(self._s.data), np.array(self._v.data))
        return (TracerArray, (self._num,), data)

    def __setstate__(self, data):
        r, s, v = data
        self._r.data[:] = r[:]
        self._s.data[:] = s[:]
        self._v.data[:] = v[:]
        self.H2D()


# ---------------------------------------------------------------------------
# coupled step and diagnostics
# ---------------------------------------------------------------------------

def system_step(Lattice lat, list swimmers, list tracers, int nsteps=1):
    """nsteps repetitions of System.Step (lbswim/System.py:9-16) in one native call."""
    cdef int nsw = len(swimmers), ntr = len(tracers), i
    cdef lbswim_swimmers* sw[64]
    cdef lbswim_tracers* tr[64]
    if nsw > 64 or ntr > 64:
        raise ValueError("at most 64 swimmer and 64 tracer arrays per system")
    for i in range(nsw):
        sw[i] = (<SwimmerArray?>swimmers[i]).handle()
    for i in range(ntr):
        tr[i] = (<TracerArray?>tracers[i]).handle()
    cdef lbswim_lattice* h = lat.handle()
    cdef int rc
    with nogil:
        rc = lbswim_system_step(h, sw, nsw, tr, ntr, nsteps)
    check(rc)


def system_autotune(Lattice lat, list swimmers, list tracers, int steps_per_trial=5):
    """Times the two launches and the fused launch's tile shapes on real
    coupled steps (the system advances by them) and keeps the fastest;
    returns (rows, strips), rows 0 for the two launches."""
    cdef int nsw = len(swimmers), ntr = len(tracers), i, rows = 0, strips = 0
    cdef lbswim_swimmers* sw[64]
    cdef lbswim_tracers* tr[64]
    if nsw > 64 or ntr > 64:
        raise ValueError("at most 64 swimmer and 64 tracer arrays per system")
    for i in range(nsw):
        sw[i] = (<SwimmerArray?>swimmers[i]).handle()
    for i in range(ntr):
        tr[i] = (<TracerArray?>tracers[i]).handle()
    check(lbswim_system_autotune(lat.handle(), sw, nsw, tr, ntr, steps_per_trial, &rows, &strips))
    return rows, strips


def device_memory():
    """(free, total) device memory in bytes."""
    cdef size_t f = 0, t = 0
    check(lbswim_device_memory(&f, &t))
    return f, t


def diag_copy(size_t count, int reps):
    """Average ms of a device copy of `count` doubles (bandwidth probe)."""
    cdef double ms = 0
    check(lbswim_diag_copy(count, reps, &ms))
    return ms


def diag_peskin(x):
    cdef cnp.ndarray[double, ndim=1, mode="c"] xa = np.ascontiguousarray(x, dtype=np.float64)
    cdef cnp.ndarray[double, ndim=1, mode="c"] out = np.zeros_like(xa)
    check(lbswim_diag_peskin(&xa[0] if xa.size else NULL, &out[0] if out.size else NULL, xa.size))
    return out


def diag_interp(size, u, r):
    cdef cnp.ndarray[double, ndim=1, mode="c"] ua = np.ascontiguousarray(u, dtype=np.float64).ravel()
    cdef cnp.ndarray[double, ndim=1, mode="c"] ra = np.ascontiguousarray(r, dtype=np.float64).ravel()
    cdef int count = ra.size // 3
    cdef cnp.ndarray[double, ndim=1, mode="c"] v = np.zeros(3 * count)
    if ua.size != 3 * size[0] * size[1] * size[2]:
        raise ValueError("u must hold 3*nx*ny*nz values")
    if count:
        check(lbswim_diag_interp(size[0], size[1], size[2], &ua[0], &ra[0], &v[0], count))
    return v.reshape(3, count)


# ---------------------------------------------------------------------------
# ensembles (no reference counterpart)
# ---------------------------------------------------------------------------

cdef class Ensemble:
    """Independent systems (lattice + swimmer and tracer arrays) advanced by
    batched launches; each member evolves exactly as it would alone.

    Ensemble(storage="f64"); add(lat, things) -> member id; step(n);
    remove(member); autotune(k); set_graph(bool).
    """
    cdef lbswim_ensemble* impl
    cdef dict _members
    cdef str _storage

    def __cinit__(self, storage="f64"):
        self.impl = NULL
        if storage not in STORAGE:
            raise ValueError("storage must be 'f64' or 'f32', not %r" % (storage,))
        check(lbswim_ensemble_create(STORAGE[storage], &self.impl))
        self._members = {}
        self._storage = storage

    def __dealloc__(self):
        if self.impl is not NULL:
            lbswim_ensemble_destroy(self.impl)
            self.impl = NULL

    def add(self, Lattice lat, things=()):
        cdef lbswim_swimmers* sw[64]
        cdef lbswim_tracers* tr[64]
        cdef int nsw = 0, ntr = 0, member = -1
        for th in things:
            if isinstance(th, SwimmerArray):
                if nsw == 64:
                    raise ValueError("at most 64 swimmer arrays per member")
                sw[nsw] = (<SwimmerArray>th).handle()
                nsw += 1
            elif isinstance(th, TracerArray):
                if ntr == 64:
                    raise ValueError("at most 64 tracer arrays per member")
                tr[ntr] = (<TracerArray>th).handle()
                ntr += 1
            else:
                raise TypeError("ensemble members take native swimmer and tracer arrays only")
        kinds = [isinstance(th, SwimmerArray) for th in things]
        if kinds != sorted(kinds, reverse=True):
            raise ValueError("list swimmer arrays before tracer arrays (the step moves them in that order)")
        check(lbswim_ensemble_add(self.impl, lat.handle(), sw, nsw, tr, ntr, &member))
        self._members[member] = (lat, tuple(things))  # keep the objects alive
        return member

    def remove(self, int member):
        check(lbswim_ensemble_remove(self.impl, member))
        del self._members[member]

    def member(self, int member):
        return self._members[member]

    @property
    def members(self):
        return dict(self._members)

    def __len__(self):
        cdef int k = 0
        check(lbswim_ensemble_size(self.impl, &k))
        return k

    def step(self, int nsteps=1):
        cdef int rc
        with nogil:
            rc = lbswim_ensemble_step(self.impl, nsteps)
        check(rc)

    def set_graph(self, bint on):
        check(lbswim_ensemble_set_graph(self.impl, on))

    def set_order(self, int order):
        check(lbswim_ensemble_set_order(self.impl, order))

    def set_force_mode(self, mode):
        """"auto" (default), "stored" or "virtual" swimmer force in the batched steps."""
        check(lbswim_ensemble_set_force_mode(self.impl, force_mode_id(mode)))

    def autotune(self, int steps_per_trial=10):
        cdef int choice = 0
        check(lbswim_ensemble_autotune(self.impl, steps_per_trial, &choice))
        return choice

    def set_work(self, int blocks_per_item):
        """Lattice blocks (256 sites) per workgroup of the batched launches."""
        check(lbswim_ensemble_set_work(self.impl, blocks_per_item))

    @property
    def tuning(self):
        """(order, blocks per work item) in use."""
        cdef int o = 0, p = 0
        check(lbswim_ensemble_tuning(self.impl, &o, &p))
        return o, p

    def set_fusion(self, int rows, int strip=32):
        """Batched fused step for the virtual-force members with full tiles:
        tiles of `rows` (8 or 4; 0 never) lattice rows x 64 sites, strips of
        `strip` planes.  Results do not depend on it."""
        check(lbswim_ensemble_set_fusion(self.impl, rows, strip))

    @property
    def fusion(self):
        """(rows, strip) of the batched fused step (rows 0: off)."""
        cdef int r = 0, st = 0
        check(lbswim_ensemble_fusion(self.impl, &r, &st))
        return r, st

    def set_lanes(self, int lanes):
        """Split the members into `lanes` groups of about equal sites whose
        batched launches run on streams of their own, side by side (1: one
        batch).  Results do not depend on it."""
        check(lbswim_ensemble_set_lanes(self.impl, lanes))

    @property
    def lanes(self):
        cdef int k = 0
        check(lbswim_ensemble_lanes(self.impl, &k))
        return k

    def fused_sites(self):
        """Lattice sites of the members the batched fused step advances."""
        cdef size_t f = 0
        check(lbswim_ensemble_fused_sites(self.impl, &f))
        return f

    def set_timing(self, on):
        """False: no timing (graphs replayed); True or an iterable of phase
        names from KERNELS ("add_forces", "step_hydro", "collide_stream",
        "calc_hydro", "stored_chain", "swimmers_move", "tracers_move"): steps
        launched eagerly with HIP events around those phases."""
        cdef unsigned mask = 0
        if on is True:
            mask = (1 << len(KERNELS)) - 1
        elif on:
            for name in on:
                mask |= 1 << KERNELS.index(name)
        check(lbswim_ensemble_set_timing(self.impl, mask))

    def kernel_times(self):
        """{phase: (total_ms, count)} recorded since set_timing."""
        cdef double ms = 0
        cdef long cnt = 0
        out = {}
        for k, name in enumerate(KERNELS):
            check(lbswim_ensemble_kernel_time(self.impl, k, &ms, &cnt))
            out[name] = (ms, cnt)
        return out

    def sites(self):
        """(sites of the virtual-force members the batched collide and
        CalcHydro launches advance, sites of stored-force members); the
        fused members' sites: fused_sites()."""
        cdef size_t v = 0, st = 0
        check(lbswim_ensemble_sites(self.impl, &v, &st))
        return v, st
