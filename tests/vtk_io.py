"""Reader of the legacy-VTK frames the reference's vtkOutputUnif writes
(ParticlesEdgesPanels2.f90:2116-2260), for tests: splits a frame into its
sections and parses the F24.18 fields (a field of '*' reads as NaN)."""
import numpy as np


def _f24(line):
    """The F24.18 fields of one record, as floats."""
    out = []
    for k in range(0, len(line), 24):
        field = line[k:k + 24].strip()
        out.append(np.nan if field.startswith("*") else float(field))
    return out


def read_frame(path):
    lines = open(path).read().split("\n")
    assert lines[0] == "# vtk DataFile Version 2.0" and lines[2] == "ASCII"
    n_active = int(lines[1].split(":")[1].split()[0])
    n_points = int(lines[4].split()[1])
    pos = 5
    points = np.array([_f24(lines[pos + j]) for j in range(n_points)]).reshape(n_points, 3)
    pos += n_points
    head = lines[pos].split()
    assert head[0] == "POLYGONS" and int(head[1]) == n_active
    pos += 1
    cells = []
    while not lines[pos].startswith("POINT_DATA"):
        cells.append([int(v) for v in lines[pos].split()])
        pos += 1
    point_data, cell_data = {}, {}
    section = point_data
    count = n_points
    while pos < len(lines) and lines[pos] != "":
        line = lines[pos]
        if line.startswith("POINT_DATA"):
            pos += 1
            continue
        if line.startswith("CELL_DATA"):
            section, count = cell_data, int(line.split()[1])
            pos += 1
            continue
        name = line.split()[1]
        assert lines[pos + 1] == "LOOKUP_TABLE default"
        section[name] = np.array([_f24(lines[pos + 2 + j])[0] for j in range(count)])
        pos += 2 + count
    return dict(n_active=n_active, points=points, cells=cells, point_data=point_data,
                cell_data=cell_data)


def rewrite(frame, path):
    """Write a parsed triangle / quadrilateral frame again with libbve's
    writer (a value read from an F24.18 field formats back to that field)."""
    from thesiscode_amd import vtk
    cells = np.array(frame["cells"])
    kind = int(cells[0, 0])
    pd, cd = frame["point_data"], frame["cell_data"]
    names = [k for k in pd if k.startswith("Tracer")]
    tr = np.column_stack([pd[k] for k in names]) if names else None
    ctr = np.column_stack([cd["TracerPanel%d" % (i + 1)] for i in range(len(names))]) \
        if names and "TracerPanel1" in cd else None
    vtk.write_frame(path, frame["points"], kind, cells[:, 1:], n_active=frame["n_active"],
                    particles_relvort=pd.get("relVort"), particles_absvort=pd.get("absVort"),
                    particles_tracer=tr, panels_tracer=ctr, panels_relvort=cd.get("relVortPanel"),
                    panels_absvort=cd.get("absVortPanel"))
