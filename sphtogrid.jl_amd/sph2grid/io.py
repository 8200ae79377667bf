"""Writers and readers of mapped images: FITS (shared/io.jl) and VTK
(shared/vtk.jl), written directly (no FITSIO / WriteVTK).

* ``write_fits_image(filename, image, par, units, snap)`` -- shared/io.jl:12-58:
  one image HDU per ``image[:, :, k]`` with the mapping parameters in the
  header (keys SNAP, CENTER_*, XMIN..ZMAX, BOXSIZE, Z_SLICE, NPIXELS,
  PIX_SIZE, UNITS).  Cards are formatted as CFITSIO formats them (the
  reference writes through FITSIO.jl), so a 2D map written here is byte for
  byte the file the reference writes.
* ``write_fits_image(filename, image, units=..., snap=...)`` without ``par``
  -- shared/io.jl:66-85 (keys SNAP, UNITS only).
* ``read_fits_image`` / ``read_allsky_fits_image`` -- shared/io.jl:110-195.
* ``save_healpix_fits`` / ``read_healpix_fits`` -- the all-sky map file of
  distributed_allsky_map (distributed_mapping/healpix.jl:79-82 calls
  Healpix.jl's ``saveToFITS``): the HEALPix binary-table convention
  (PIXTYPE, ORDERING = RING, NSIDE, one column of 1024 values per row).
* ``write_vtk_image`` -- shared/vtk.jl:10-21: a VTK XML RectilinearGrid
  (``.vtr``) on the pixel centres of ``get_map_grid_3D``
  (shared/reconstruct_grid.jl:27-50), point data ``image_name``, field data
  ``Units`` and ``Snap``; raw appended binary with UInt64 block headers.
"""
from __future__ import annotations

import numpy as np

from .parameters import MappingParameters, mappingParameters

_BLOCK = 2880
_FITS_COMMENTS = (
    "FITS (Flexible Image Transport System) format is defined in 'Astronomy",
    "and Astrophysics', volume 376, page 359; bibcode: 2001A&A...376..359H",
)


# --------------------------------------------------------------------------- cards
def _fmt_value(v) -> str:
    """Value field of a header card as CFITSIO writes it (right-justified
    numbers in columns 11-30, quoted strings left-justified)."""
    if isinstance(v, (bool, np.bool_)):
        return ("T" if v else "F").rjust(20)
    if isinstance(v, (int, np.integer)):
        return str(int(v)).rjust(20)
    if isinstance(v, (float, np.floating)):
        s = "%.15G" % float(v)
        if "." not in s and "E" in s:
            s = "%.1E" % float(v)
        if "." not in s and "E" not in s and "N" not in s and "I" not in s:
            s += "."
        return s.rjust(20)
    s = str(v).replace("'", "''")
    return ("'" + s.ljust(8) + "'").ljust(20)


def _card(key: str, value=None, comment: str = "") -> bytes:
    if key == "COMMENT":
        c = "COMMENT   " + comment
    elif key == "END":
        c = "END"
    else:
        c = key.ljust(8)[:8] + "= " + _fmt_value(value)
        if comment:
            c += " / " + comment
    return c[:80].ljust(80).encode("ascii")


def _pad(b: bytes, fill: bytes) -> bytes:
    r = len(b) % _BLOCK
    return b if r == 0 else b + fill * (_BLOCK - r)


def _image_hdu(data: np.ndarray, keys, primary: bool) -> bytes:
    """One image HDU; ``data`` is a Julia-ordered array (first axis fastest)."""
    data = np.asarray(data, dtype=np.float64)
    cards = [_card("SIMPLE", True, "file does conform to FITS standard")] if primary else \
        [_card("XTENSION", "IMAGE", "IMAGE extension")]
    cards.append(_card("BITPIX", -64, "number of bits per data pixel"))
    cards.append(_card("NAXIS", data.ndim, "number of data axes"))
    for k, n in enumerate(data.shape):
        cards.append(_card(f"NAXIS{k + 1}", int(n), f"length of data axis {k + 1}"))
    if primary:
        cards.append(_card("EXTEND", True, "FITS dataset may contain extensions"))
        cards += [_card("COMMENT", comment=c) for c in _FITS_COMMENTS]
    else:
        cards.append(_card("PCOUNT", 0, "required keyword; must = 0"))
        cards.append(_card("GCOUNT", 1, "required keyword; must = 1"))
    cards += [_card(k, v, c) for k, v, c in keys]
    cards.append(_card("END"))
    head = _pad(b"".join(cards), b" ")
    body = np.asfortranarray(data).ravel(order="F").astype(">f8").tobytes()
    return head + _pad(body, b"\0")


def _param_keys(par: MappingParameters, units: str, snap: int):
    """shared/io.jl:18-45"""
    z_slice = abs(float(par.z_lim[0])) + abs(float(par.z_lim[1]))
    return [
        ("SNAP", int(snap), "snapshot number"),
        ("CENTER_X", float(par.center[0]), "image center (x)"),
        ("CENTER_Y", float(par.center[1]), "image center (y)"),
        ("CENTER_Z", float(par.center[2]), "image center (z)"),
        ("XMIN", float(par.x_lim[0]), "x limit left"),
        ("XMAX", float(par.x_lim[1]), "x limit right"),
        ("YMIN", float(par.y_lim[0]), "y limit left"),
        ("YMAX", float(par.y_lim[1]), "y limit right"),
        ("ZMIN", float(par.z_lim[0]), "z limit left"),
        ("ZMAX", float(par.z_lim[1]), "z limit right"),
        ("BOXSIZE", float(par.boxsize), "size of the image"),
        ("Z_SLICE", z_slice, "depth of the integrated slice"),
        ("NPIXELS", int(np.max(par.Npixels)), "image resolution"),
        ("PIX_SIZE", float(par.pixelSideLength), "pixel size"),
        ("UNITS", units, "units of the image"),
    ]


def write_fits_image(filename: str, image, par: MappingParameters | None = None, *,
                     units: str = "[i.u.]", snap: int = 0) -> None:
    """shared/io.jl:12-58 (with ``par``) and :66-85 (without)."""
    image = np.asarray(image, dtype=np.float64)
    if par is None:
        blob = _image_hdu(image, [("SNAP", int(snap), "snapshot number"),
                                  ("UNITS", units, "units of the image")], True)
    else:
        keys = _param_keys(par, units, snap)
        planes = [image] if image.ndim < 3 else [image[:, :, k] for k in range(image.shape[2])]
        blob = b"".join(_image_hdu(p, keys, k == 0) for k, p in enumerate(planes))
    with open(filename, "wb") as f:
        f.write(blob)


# --------------------------------------------------------------------------- reading
def _parse_value(s: str):
    s = s.strip()
    if s.startswith("'"):
        end = s.rfind("'")
        return s[1:end].replace("''", "'").rstrip()
    if s in ("T", "F"):
        return s == "T"
    try:
        return int(s)
    except ValueError:
        return float(s.replace("D", "E"))


def _read_header(buf: bytes, pos: int):
    hdr = {}
    while True:
        block = buf[pos:pos + _BLOCK]
        if len(block) < _BLOCK:
            raise ValueError("truncated FITS header")
        pos += _BLOCK
        done = False
        for k in range(0, _BLOCK, 80):
            card = block[k:k + 80].decode("ascii")
            key = card[:8].strip()
            if key == "END":
                done = True
                break
            if card[8:10] == "= ":
                val = card[10:]
                if val.lstrip().startswith("'"):
                    q = val.index("'")
                    j = q + 1
                    while True:
                        j = val.index("'", j)
                        if j + 1 < len(val) and val[j + 1] == "'":
                            j += 2
                            continue
                        break
                    hdr[key] = _parse_value(val[q:j + 1])
                else:
                    hdr[key] = _parse_value(val.split("/")[0])
        if done:
            return hdr, pos


def _read_hdus(filename: str):
    buf = open(filename, "rb").read()
    pos, hdus = 0, []
    while pos < len(buf):
        hdr, pos = _read_header(buf, pos)
        naxis = hdr.get("NAXIS", 0)
        shape = [hdr[f"NAXIS{k + 1}"] for k in range(naxis)]
        bitpix = hdr["BITPIX"]
        nbytes = abs(bitpix) // 8 * int(np.prod(shape)) if naxis else 0
        if hdr.get("XTENSION") == "BINTABLE":
            nbytes = hdr["NAXIS1"] * hdr["NAXIS2"] + hdr.get("PCOUNT", 0)
        raw = buf[pos:pos + nbytes]
        pos += -(-nbytes // _BLOCK) * _BLOCK
        hdus.append((hdr, raw, shape))
    return hdus


def _image_of(hdr, raw, shape):
    dt = {-64: ">f8", -32: ">f4", 8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8"}[hdr["BITPIX"]]
    return np.frombuffer(raw, dtype=dt).astype(np.float64).reshape(shape, order="F")


def read_fits_image(filename: str, Nimage: int = 1, verbose: bool = False):
    """shared/io.jl:110-150: (image, mappingParameters, snap, units)."""
    hdr, raw, shape = _read_hdus(filename)[Nimage - 1]
    image = _image_of(hdr, raw, shape)
    par = mappingParameters(x_lim=[hdr["XMIN"], hdr["XMAX"]], y_lim=[hdr["YMIN"], hdr["YMAX"]],
                            z_lim=[hdr["ZMIN"], hdr["ZMAX"]], Npixels=int(hdr["NPIXELS"]),
                            boxsize=hdr["BOXSIZE"])
    return image, par, hdr["SNAP"], hdr["UNITS"]


def read_allsky_fits_image(filename: str, verbose: bool = False):
    """shared/io.jl:165-195: (image, snap, units)."""
    hdr, raw, shape = _read_hdus(filename)[0]
    return _image_of(hdr, raw, shape), hdr["SNAP"], hdr["UNITS"]


# --------------------------------------------------------------------------- HEALPix
def save_healpix_fits(filename: str, pixels, nside: int, *, column: str = "TEMPERATURE") -> None:
    """All-sky map in the HEALPix FITS convention (RING ordering): an empty
    primary HDU and a binary table with one column of 1024 Float64 per row."""
    pixels = np.asarray(pixels, dtype=np.float64).ravel()
    npix = 12 * nside * nside
    if pixels.size != npix:
        raise ValueError(f"map has {pixels.size} pixels, Nside {nside} needs {npix}")
    per_row = 1024 if npix >= 1024 else npix
    nrows = npix // per_row
    prim = [_card("SIMPLE", True, "file does conform to FITS standard"),
            _card("BITPIX", 8, "number of bits per data pixel"),
            _card("NAXIS", 0, "number of data axes"),
            _card("EXTEND", True, "FITS dataset may contain extensions")]
    prim += [_card("COMMENT", comment=c) for c in _FITS_COMMENTS]
    prim.append(_card("END"))
    tab = [_card("XTENSION", "BINTABLE", "binary table extension"),
           _card("BITPIX", 8, "8-bit bytes"),
           _card("NAXIS", 2, "2-dimensional binary table"),
           _card("NAXIS1", 8 * per_row, "width of table in bytes"),
           _card("NAXIS2", nrows, "number of rows in table"),
           _card("PCOUNT", 0, "size of special data area"),
           _card("GCOUNT", 1, "one data group (required keyword)"),
           _card("TFIELDS", 1, "number of fields in each row"),
           _card("TTYPE1", column, "label for field   1"),
           _card("TFORM1", f"{per_row}D", "data format of field: 8-byte DOUBLE"),
           _card("PIXTYPE", "HEALPIX", "HEALPIX pixelisation"),
           _card("ORDERING", "RING", "Pixel ordering scheme, either RING or NESTED"),
           _card("NSIDE", int(nside), "Resolution parameter for HEALPIX"),
           _card("FIRSTPIX", 0, "First pixel # (0 based)"),
           _card("LASTPIX", npix - 1, "Last pixel # (0 based)"),
           _card("INDXSCHM", "IMPLICIT", "Indexing: IMPLICIT or EXPLICIT"),
           _card("END")]
    blob = _pad(b"".join(prim), b" ") + _pad(b"".join(tab), b" ") + \
        _pad(pixels.astype(">f8").tobytes(), b"\0")
    with open(filename, "wb") as f:
        f.write(blob)


def read_healpix_fits(filename: str):
    """(pixels in RING order, Nside) of a file written by save_healpix_fits."""
    hdr, raw, _ = [h for h in _read_hdus(filename) if h[0].get("XTENSION") == "BINTABLE"][0]
    return np.frombuffer(raw, dtype=">f8").astype(np.float64), int(hdr["NSIDE"])


# --------------------------------------------------------------------------- VTK
def get_map_grid_3D(par: MappingParameters):
    """shared/reconstruct_grid.jl:27-50: pixel centres along x, y, z."""
    n = [int(v) for v in par.Npixels]
    return tuple(float(lim[0]) + (np.arange(1, k + 1) - 0.5) * par.pixelSideLength
                 for lim, k in zip((par.x_lim, par.y_lim, par.z_lim), n))


def write_vtk_image(filename: str, image, image_name: str, par: MappingParameters, *,
                    units: str = "[i.u.]", snap: int = 0) -> str:
    """shared/vtk.jl:10-21.  ``image`` is the 3D map as returned by the 3D
    sphMapping (Julia order, first axis fastest).  Writes ``filename.vtr``
    and returns the path."""
    x, y, z = get_map_grid_3D(par)
    image = np.asarray(image, dtype=np.float64)
    if image.shape != (x.size, y.size, z.size):
        raise ValueError(f"image shape {image.shape} does not match the grid "
                         f"{(x.size, y.size, z.size)}")
    path = filename if filename.endswith(".vtr") else filename + ".vtr"
    blocks, offsets, off = [], [], 0

    def add(arr: np.ndarray):
        nonlocal off
        data = np.ascontiguousarray(arr).tobytes()
        blocks.append(np.uint64(len(data)).tobytes() + data)
        offsets.append(off)
        off += 8 + len(data)

    add(x)
    add(y)
    add(z)
    add(image.ravel(order="F"))
    ub = np.frombuffer(units.encode("utf-8"), dtype=np.uint8)
    add(np.append(ub, np.uint8(0)))
    add(np.array([int(snap)], dtype=np.int64))
    ext = f"0 {x.size - 1} 0 {y.size - 1} 0 {z.size - 1}"
    xml = (
        '<?xml version="1.0" encoding="utf-8"?>\n'
        '<VTKFile type="RectilinearGrid" version="1.0" byte_order="LittleEndian" '
        'header_type="UInt64">\n'
        f'  <RectilinearGrid WholeExtent="{ext}">\n'
        '    <FieldData>\n'
        f'      <DataArray type="UInt8" Name="Units" NumberOfTuples="{ub.size + 1}" '
        f'format="appended" offset="{offsets[4]}"/>\n'
        f'      <DataArray type="Int64" Name="Snap" NumberOfTuples="1" format="appended" '
        f'offset="{offsets[5]}"/>\n'
        '    </FieldData>\n'
        f'    <Piece Extent="{ext}">\n'
        '      <Coordinates>\n'
        f'        <DataArray type="Float64" Name="x" format="appended" offset="{offsets[0]}"/>\n'
        f'        <DataArray type="Float64" Name="y" format="appended" offset="{offsets[1]}"/>\n'
        f'        <DataArray type="Float64" Name="z" format="appended" offset="{offsets[2]}"/>\n'
        '      </Coordinates>\n'
        f'      <PointData Scalars="{image_name}">\n'
        f'        <DataArray type="Float64" Name="{image_name}" format="appended" '
        f'offset="{offsets[3]}"/>\n'
        '      </PointData>\n'
        '    </Piece>\n'
        '  </RectilinearGrid>\n'
        '  <AppendedData encoding="raw">\n   _')
    with open(path, "wb") as f:
        f.write(xml.encode("utf-8"))
        for b in blocks:
            f.write(b)
        f.write(b"\n  </AppendedData>\n</VTKFile>\n")
    return path


def read_vtk_image(path: str):
    """Inverse of write_vtk_image: (image, (x, y, z), units, snap)."""
    import xml.etree.ElementTree as ET
    raw = open(path, "rb").read()
    start = raw.index(b"<AppendedData")
    us = raw.index(b"_", start) + 1
    end = raw.rindex(b"</AppendedData>")
    app = raw[us:end]
    head = raw[:start].decode("utf-8") + "</VTKFile>"
    root = ET.fromstring(head)
    grid = root.find("RectilinearGrid")
    ext = [int(v) for v in grid.get("WholeExtent").split()]
    shape = (ext[1] + 1, ext[3] + 1, ext[5] + 1)

    def get(da):
        off = int(da.get("offset"))
        n = int(np.frombuffer(app[off:off + 8], dtype=np.uint64)[0])
        dt = {"Float64": "<f8", "Int64": "<i8", "UInt8": "u1"}[da.get("type")]
        return np.frombuffer(app[off + 8:off + 8 + n], dtype=dt)

    coords = [get(d) for d in grid.find("Piece/Coordinates")]
    pd = grid.find("Piece/PointData/DataArray")
    image = get(pd).reshape(shape, order="F")
    fd = {d.get("Name"): get(d) for d in grid.find("FieldData")}
    units = bytes(fd["Units"][:-1]).decode("utf-8")
    return image, tuple(coords), units, int(fd["Snap"][0])

