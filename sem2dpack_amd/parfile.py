"""Reader for the SEM2DPACK input file ``Par.inp``.

The input file is a sequence of Fortran NAMELIST groups (``&NAME key=value, ... /``)
that the reference reads with ``rewind`` + ``read(iin, NAME)`` in each module's
``*_read`` routine (e.g. ``SRC/input.f90:101-169``, ``SRC/time.f90:122-315``,
``SRC/mesh_cartesian.f90:98-113``).  Some groups are followed by free-format data
lines that the reference reads list-directed from the same unit, e.g. the values
of ``&DIST_PWCONR`` (``SRC/distribution_pwconr.f90``).  Text outside groups is
ignored, exactly as a namelist read skips it.

``read_par(path)`` returns a :class:`ParFile` holding the groups in file order;
each :class:`Group` keeps its (lower-case) keys, its values converted to Python
(int, float, bool, str or lists of them) and the raw data lines that follow it.
"""
import re

__all__ = ["Group", "ParFile", "read_par", "parse_par"]


class Group:
    def __init__(self, name, values, data_lines, index):
        self.name = name.upper()
        self.values = values
        self.data_lines = data_lines
        self.index = index

    def get(self, key, default=None):
        return self.values.get(key.lower(), default)

    def __contains__(self, key):
        return key.lower() in self.values

    def __repr__(self):
        return f"Group({self.name}, {self.values})"


class ParFile:
    def __init__(self, groups):
        self.groups = groups

    def all(self, name):
        name = name.upper()
        return [g for g in self.groups if g.name == name]

    def first(self, name):
        g = self.all(name)
        return g[0] if g else None

    def after(self, group, name):
        """First group called ``name`` that follows ``group`` in the file."""
        name = name.upper()
        for g in self.groups[group.index + 1:]:
            if g.name == name:
                return g
        return None

    def next_groups(self, group):
        return self.groups[group.index + 1:]


_TOKEN = re.compile(r"""\s*(?:('(?:[^']|'')*')|("(?:[^"]|"")*")|([^,\s=/]+)|(,)|(=)|(/))""")


def _convert(tok):
    if tok[0] in "'\"":
        q = tok[0]
        return tok[1:-1].replace(q + q, q)
    low = tok.lower()
    if low in (".true.", "t", ".t.", "true"):
        return True
    if low in (".false.", "f", ".f.", "false"):
        return False
    try:
        return int(tok)
    except ValueError:
        pass
    try:
        return float(low.replace("d", "e"))
    except ValueError:
        return tok


def _parse_body(body):
    """Parse 'key=v1,v2, key2=...' into a dict (values: scalar or list)."""
    toks = []
    pos = 0
    while pos < len(body):
        m = _TOKEN.match(body, pos)
        if not m or m.end() == pos:
            break
        pos = m.end()
        for i, t in enumerate(m.groups()):
            if t is not None:
                toks.append((i, t))
                break
    values = {}
    key = None
    cur = []
    i = 0
    while i < len(toks):
        kind, t = toks[i]
        if kind == 2 and i + 1 < len(toks) and toks[i + 1][0] == 4:
            if key is not None:
                values[key] = cur[0] if len(cur) == 1 else cur
            key = t.lower()
            cur = []
            i += 2
            continue
        if kind in (0, 1, 2):
            m = re.match(r"^(\d+)\*(.+)$", t) if kind == 2 else None
            if m:
                cur.extend([_convert(m.group(2))] * int(m.group(1)))
            else:
                cur.append(_convert(t))
        i += 1
    if key is not None:
        values[key] = cur[0] if len(cur) == 1 else cur
    return values


def _strip_comment(line):
    out, q = [], None
    for ch in line:
        if q:
            out.append(ch)
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
            out.append(ch)
        elif ch in "#!":
            break
        else:
            out.append(ch)
    return "".join(out)


def parse_par(text):
    lines = text.splitlines()
    groups = []
    i = 0
    n = len(lines)
    while i < n:
        line = _strip_comment(lines[i])
        m = re.search(r"&(\w+)", line)
        if not m:
            i += 1
            continue
        name = m.group(1)
        body = line[m.end():]
        # accumulate until the terminating '/' outside quotes
        while True:
            if _terminated(body):
                break
            i += 1
            if i >= n:
                break
            body += " " + _strip_comment(lines[i])
        body, rest = _split_term(body)
        values = _parse_body(body)
        data = []
        if rest.strip():
            data.append(rest.strip())
        i += 1
        while i < n and not re.search(r"&\w+", _strip_comment(lines[i])):
            s = _strip_comment(lines[i]).strip()
            if s:
                data.append(s)
            i += 1
        groups.append(Group(name, values, data, len(groups)))
    return ParFile(groups)


def _terminated(body):
    q = None
    for ch in body:
        if q:
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
        elif ch == "/":
            return True
    return False


def _split_term(body):
    q = None
    for k, ch in enumerate(body):
        if q:
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
        elif ch == "/":
            return body[:k], body[k + 1:]
    return body, ""


def read_par(path):
    with open(path) as fh:
        return parse_par(fh.read())
