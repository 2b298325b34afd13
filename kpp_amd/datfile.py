"""KPP's ROOT.dat output (util/util.c:17-43: InitSaveData, SaveData,
CloseSaveData) for cells of a batch.

SaveData writes one line per call: TIME/3600 as "%6.1f " and then, for every
#LOOKAT species, C[species]/CFACTOR as "%24.16e ", then a newline.  The
reference writes one file for its one box; here each selected cell of a
batch gets its own file, named by `path` with "{cell}" replaced by the cell
index (a path without "{cell}" takes a single cell).
"""
from __future__ import annotations

import numpy as np


class SaveData:
    def __init__(self, path: str, species: list[str], lookat: list[str], cells):
        cells = [int(c) for c in np.atleast_1d(cells)]
        if "{cell}" not in path and len(cells) != 1:
            raise ValueError("a path without {cell} takes exactly one cell")
        index = {s: k for k, s in enumerate(species)}
        missing = [s for s in lookat if s not in index]
        if missing:
            raise KeyError(f"#LOOKAT species not in the mechanism: {missing}")
        self.idx = [index[s] for s in lookat]
        self.cells = cells
        # InitSaveData: the file is created (truncated) when saving starts
        self.files = [open(path.replace("{cell}", str(c)), "w") for c in cells]

    def save(self, time, C, cfactor):
        """C: host [NSPEC, ncells]; time and cfactor: scalars or per-cell arrays"""
        C = np.asarray(C)
        t = np.broadcast_to(np.asarray(time, np.float64), (C.shape[1],))
        cf = np.broadcast_to(np.asarray(cfactor, np.float64), (C.shape[1],))
        for f, c in zip(self.files, self.cells):
            line = "%6.1f " % (t[c] / 3600.0)
            line += "".join("%24.16e " % (C[k, c] / cf[c]) for k in self.idx)
            f.write(line + "\n")

    def close(self):
        for f in self.files:
            f.close()
        self.files = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
