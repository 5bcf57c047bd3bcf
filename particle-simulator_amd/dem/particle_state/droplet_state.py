"""Single droplet-system state: positions (N,3), radii (N,), time.

Drop-in for the reference's ``DropletState``
(dem/particle_state/droplet_state.py:7-100): construction from arrays or
from a ``.npz`` / ``.xlsx`` file, ``copy``, ``save``/``load`` (npz keys
positions, radii, time) and xlsx export/import with columns x, y, z, r (time
is not stored in xlsx).  The xlsx path is a small self-contained
Office-Open-XML reader/writer (zip + XML), because the reference's pandas /
openpyxl route is not available in this image; it reads files written by
pandas/openpyxl (numeric cells, first sheet, header row) and writes files
they can read.
"""

from __future__ import annotations

import re
import zipfile
from xml.sax.saxutils import escape

import numpy as np

from .particle_state_base import ParticleState


def _col_letter(k: int) -> str:
    s = ""
    k += 1
    while k:
        k, rem = divmod(k - 1, 26)
        s = chr(65 + rem) + s
    return s


def write_xlsx(filename: str, columns: dict[str, np.ndarray]) -> None:
    """Write numeric columns (header row + values) to the first sheet."""
    names = list(columns)
    data = [np.asarray(columns[c], dtype=np.float64) for c in names]
    nrows = len(data[0]) if data else 0
    rows = []
    hdr = "".join(
        f'<c r="{_col_letter(k)}1" t="inlineStr"><is><t>{escape(n)}</t></is></c>'
        for k, n in enumerate(names)
    )
    rows.append(f'<row r="1">{hdr}</row>')
    for i in range(nrows):
        cells = "".join(
            f'<c r="{_col_letter(k)}{i + 2}"><v>{repr(float(data[k][i]))}</v></c>'
            for k in range(len(names))
        )
        rows.append(f'<row r="{i + 2}">{cells}</row>')
    sheet = (
        '<?xml version="1.0" encoding="UTF-8" standalone="yes"?>'
        '<worksheet xmlns="http://schemas.openxmlformats.org/spreadsheetml/2006/main">'
        f'<sheetData>{"".join(rows)}</sheetData></worksheet>'
    )
    ct = (
        '<?xml version="1.0" encoding="UTF-8" standalone="yes"?>'
        '<Types xmlns="http://schemas.openxmlformats.org/package/2006/content-types">'
        '<Default Extension="rels" ContentType="application/vnd.openxmlformats-package.relationships+xml"/>'
        '<Default Extension="xml" ContentType="application/xml"/>'
        '<Override PartName="/xl/workbook.xml" ContentType="application/vnd.openxmlformats-officedocument.spreadsheetml.sheet.main+xml"/>'
        '<Override PartName="/xl/worksheets/sheet1.xml" ContentType="application/vnd.openxmlformats-officedocument.spreadsheetml.worksheet+xml"/>'
        "</Types>"
    )
    rels = (
        '<?xml version="1.0" encoding="UTF-8" standalone="yes"?>'
        '<Relationships xmlns="http://schemas.openxmlformats.org/package/2006/relationships">'
        '<Relationship Id="rId1" Type="http://schemas.openxmlformats.org/officeDocument/2006/relationships/officeDocument" Target="xl/workbook.xml"/>'
        "</Relationships>"
    )
    wb = (
        '<?xml version="1.0" encoding="UTF-8" standalone="yes"?>'
        '<workbook xmlns="http://schemas.openxmlformats.org/spreadsheetml/2006/main" '
        'xmlns:r="http://schemas.openxmlformats.org/officeDocument/2006/relationships">'
        '<sheets><sheet name="Sheet1" sheetId="1" r:id="rId1"/></sheets></workbook>'
    )
    wb_rels = (
        '<?xml version="1.0" encoding="UTF-8" standalone="yes"?>'
        '<Relationships xmlns="http://schemas.openxmlformats.org/package/2006/relationships">'
        '<Relationship Id="rId1" Type="http://schemas.openxmlformats.org/officeDocument/2006/relationships/worksheet" Target="worksheets/sheet1.xml"/>'
        "</Relationships>"
    )
    with zipfile.ZipFile(filename, "w", zipfile.ZIP_DEFLATED) as z:
        z.writestr("[Content_Types].xml", ct)
        z.writestr("_rels/.rels", rels)
        z.writestr("xl/workbook.xml", wb)
        z.writestr("xl/_rels/workbook.xml.rels", wb_rels)
        z.writestr("xl/worksheets/sheet1.xml", sheet)


_CELL = re.compile(r"<c\b([^>]*?)(?:/>|>(.*?)</c>)", re.S)
_REF = re.compile(r'\br="([A-Z]+)(\d+)"')
_TYPE = re.compile(r'\bt="([^"]+)"')
_V = re.compile(r"<v>(.*?)</v>", re.S)
_T = re.compile(r"<t[^>]*>(.*?)</t>", re.S)


def _col_index(letters: str) -> int:
    k = 0
    for ch in letters:
        k = k * 26 + (ord(ch) - 64)
    return k - 1


def read_xlsx(filename: str) -> dict[str, np.ndarray]:
    """Read the first sheet: header row -> float64 columns."""
    with zipfile.ZipFile(filename) as z:
        names = z.namelist()
        shared: list[str] = []
        if "xl/sharedStrings.xml" in names:
            sx = z.read("xl/sharedStrings.xml").decode("utf-8")
            for si in re.findall(r"<si>(.*?)</si>", sx, re.S):
                shared.append("".join(_T.findall(si)))
        sheet = "xl/worksheets/sheet1.xml"
        if sheet not in names:
            sheet = sorted(n for n in names if n.startswith("xl/worksheets/sheet"))[0]
        xml = z.read(sheet).decode("utf-8")
    table: dict[int, dict[int, object]] = {}
    for attrs, body in _CELL.findall(xml):
        ref = _REF.search(attrs)
        if not ref:
            continue
        col, row = _col_index(ref.group(1)), int(ref.group(2))
        typ = _TYPE.search(attrs)
        typ = typ.group(1) if typ else "n"
        body = body or ""
        if typ == "s":
            val = shared[int(_V.search(body).group(1))]
        elif typ == "inlineStr":
            val = "".join(_T.findall(body))
        else:
            v = _V.search(body)
            val = float(v.group(1)) if v else np.nan
        table.setdefault(row, {})[col] = val
    rows = sorted(table)
    header = table[rows[0]]
    out = {}
    for col, name in header.items():
        out[str(name)] = np.array([float(table[r].get(col, np.nan)) for r in rows[1:]])
    return out


class DropletState(ParticleState):
    def __init__(self, positions=None, radii=None, time=0, filename=None):
        self.positions = None
        self.radii = None
        self.time = 0
        if positions is not None and radii is not None:
            self.positions = np.array(positions)
            self.radii = np.array(radii)
            self.time = time
        elif filename is not None:
            if filename.endswith(".xlsx"):
                self.import_from_xlsx(filename)
            elif filename.endswith(".npz"):
                self.load(filename)
            else:
                raise ValueError("file must have a .xlsx or .npz extension")
        else:
            raise ValueError("pass either positions and radii, or a file name")

    def copy(self):
        return DropletState(self.positions.copy(), self.radii.copy(), self.time)

    @classmethod
    def adopt(cls, positions, radii, time=0):
        """A state holding ``positions`` and ``radii`` themselves (no copy), for
        arrays that nothing else refers to."""
        st = cls.__new__(cls)
        st.positions, st.radii, st.time = positions, radii, time
        return st

    def save(self, filename):
        np.savez(filename, positions=self.positions, radii=self.radii, time=self.time)

    def load(self, filename):
        data = np.load(filename)
        self.positions = data["positions"]
        self.radii = data["radii"]
        self.time = float(data["time"])

    def export_to_xlsx(self, filename):
        write_xlsx(
            filename,
            {
                "x": self.positions[:, 0],
                "y": self.positions[:, 1],
                "z": self.positions[:, 2],
                "r": self.radii,
            },
        )

    def import_from_xlsx(self, filename):
        cols = read_xlsx(filename)
        self.positions = np.stack([cols["x"], cols["y"], cols["z"]], axis=1)
        self.radii = np.array(cols["r"])
        self.time = 0

    def __repr__(self):
        return f"<DropletState: time={self.time}, num_particles={len(self.radii)}>"
