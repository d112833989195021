import numpy as np


def rel_fro_lower(x, ref):
    """||tril(x - ref)||_F / ||tril(ref)||_F over the lower triangle incl. diagonal."""
    d = np.tril(x - ref)
    return float(np.linalg.norm(d) / np.linalg.norm(np.tril(ref)))


def rel_fro(x, ref):
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


def require_dump_io():
    """Load the dump I/O library: a failure where HDF5 is installed (build()
    builds the library there), a skip that says so on a machine without HDF5."""
    import pytest
    from hybridhsdla_amd import build, dump as dump_mod
    try:
        return dump_mod.load_io()
    except (RuntimeError, OSError) as e:
        if build.hdf5_available():
            pytest.fail(f"HDF5 is installed under {build.HDF5_ROOT} but the dump I/O library failed: {e}")
        pytest.skip(f"no HDF5 under {build.HDF5_ROOT}: the dump I/O library is not built ({e})")


def open_dump_or_fail(path):
    """Open a Fleur dump with the dump I/O library.  Where HDF5 is installed
    the library is built with the rest (build()), so a reader that does not
    load is a failure; only a machine without HDF5 skips, and says so."""
    import pytest
    from hybridhsdla_amd import build, dump as dump_mod
    try:
        return dump_mod.FleurDump(path)
    except (RuntimeError, OSError) as e:
        if build.hdf5_available():
            pytest.fail(f"HDF5 is installed under {build.HDF5_ROOT} but the dump reader failed: {e}")
        pytest.skip(f"no HDF5 under {build.HDF5_ROOT}: the dump reader is not built ({e})")


def stratified_sample(ng, block=128, seed=1):
    """One index drawn from every block of `block` consecutive indices."""
    rng = np.random.default_rng(seed)
    return np.array([b0 + int(rng.integers(0, min(block, ng - b0))) for b0 in range(0, ng, block)], dtype=np.int64)


def driver_or_fail():
    """Path of the built C++ driver program (hsgen_driver).  Where HDF5 is
    installed build() builds it, so a missing program is a failure; only a
    machine without HDF5 skips, and says so."""
    import pytest
    from hybridhsdla_amd import build
    if build.DRIVER.exists():
        return build.DRIVER
    if build.hdf5_available():
        pytest.fail(f"HDF5 is installed under {build.HDF5_ROOT} but {build.DRIVER.name} is not built: run build()")
    pytest.skip(f"no HDF5 under {build.HDF5_ROOT}: the driver program is not built")
