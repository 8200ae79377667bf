"""FITS / VTK writers and readers (shared/io.jl, shared/vtk.jl).  CPU only.

Pinned: the 2D FITS writer reproduces, byte for byte, the reference's own
``test/sedov_rho_reference.fits`` (written by its write_fits_image through
FITSIO.jl / CFITSIO) from that file's pixels and parameters; the reader
recovers them.  Parity unpinned: the exact card text CFITSIO gives floats
with exponents and the header of image extensions (no fixture holds them),
Healpix.jl's ``saveToFITS`` header beyond the HEALPix convention keys, and
WriteVTK.jl's exact bytes (it compresses with zlib; this writer stores raw)."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "sedov_rho_reference.fits")


def test_fits_reader_on_reference_file(s2g):
    img, par, snap, units = s2g.read_fits_image(GOLDEN)
    assert img.shape == (256, 256)
    assert snap == 50 and units == "g/cm^2"
    np.testing.assert_allclose([img.min(), img.max()], [0.021467919058626866, 0.04757891903153075])
    assert list(par.x_lim) == [0.3, 5.7] and list(par.center) == [3.0, 3.0, 3.0]
    assert par.boxsize == 6.0 and par.pixelSideLength == 0.02109375
    assert int(par.Npixels[0]) == 256


def test_fits_writer_is_byte_identical_to_reference(s2g, tmp_path):
    img, par, snap, units = s2g.read_fits_image(GOLDEN)
    out = tmp_path / "rho.fits"
    s2g.write_fits_image(str(out), img, par, units=units, snap=snap)
    assert out.read_bytes() == open(GOLDEN, "rb").read()


def test_fits_multi_image_and_allsky_round_trip(s2g, tmp_path):
    par = s2g.mappingParameters(x_lim=[-1.0, 2.0], y_lim=[-1.0, 2.0], z_lim=[-0.5, 0.5],
                                Npixels=32)
    r = np.random.default_rng(3)
    img = r.standard_normal((32, 32, 3))
    img[0, 0, 1] = np.nan
    f = str(tmp_path / "q.fits")
    s2g.write_fits_image(f, img, par, units="K", snap=7)
    assert os.path.getsize(f) % 2880 == 0
    for k in range(3):
        got, p2, snap, units = s2g.read_fits_image(f, k + 1)
        np.testing.assert_array_equal(got, img[:, :, k])
        assert snap == 7 and units == "K" and list(p2.z_lim) == [-0.5, 0.5]
    sky = r.random(12 * 8 * 8)
    f2 = str(tmp_path / "sky.fits")
    s2g.write_fits_image(f2, sky, units="[i.u.]", snap=3)
    got, snap, units = s2g.read_allsky_fits_image(f2)
    np.testing.assert_array_equal(got, sky)
    assert snap == 3 and units == "[i.u.]"


@pytest.mark.parametrize("nside", [1, 4, 16])
def test_healpix_fits_round_trip(s2g, tmp_path, nside):
    sky = np.random.default_rng(nside).random(12 * nside * nside)
    f = str(tmp_path / "sky.fits")
    s2g.save_healpix_fits(f, sky, nside)
    assert os.path.getsize(f) % 2880 == 0
    got, ns = s2g.read_healpix_fits(f)
    assert ns == nside
    np.testing.assert_array_equal(got, sky)
    head = open(f, "rb").read(2 * 2880 + 2880).decode("ascii", "replace")
    for key in ("PIXTYPE = 'HEALPIX '", "ORDERING= 'RING    '", "INDXSCHM= 'IMPLICIT'"):
        assert key in head
    with pytest.raises(ValueError):
        s2g.save_healpix_fits(f, sky[:-1], nside)


def test_vtk_round_trip_and_grid(s2g, tmp_path):
    par = s2g.mappingParameters(x_lim=[0.0, 4.0], y_lim=[0.0, 4.0], z_lim=[0.0, 4.0], Npixels=8)
    img = np.random.default_rng(1).random((8, 8, 8))
    path = s2g.write_vtk_image(str(tmp_path / "cube"), img, "map", par, units="g/cm^3", snap=11)
    assert path.endswith(".vtr")
    got, (x, y, z), units, snap = s2g.read_vtk_image(path)
    np.testing.assert_array_equal(got, img)
    # reconstruct_grid.jl:27-50: lim[1] + (i - 1/2) * pixelSideLength
    np.testing.assert_allclose(x, 0.25 + 0.5 * np.arange(8))
    np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(y, z)
    assert units == "g/cm^3" and snap == 11
    with pytest.raises(ValueError):
        s2g.write_vtk_image(str(tmp_path / "bad"), img[:4], "map", par)
