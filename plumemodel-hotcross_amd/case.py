"""Case set-up from files, as the reference takes it:

- the stratification table that INITIALIZE_PROFILES reads from unit 1010
  (initialize_profiles80.f:29-44): five header lines, then rows of
  (3F15.7) z, T, S; and
- the case file, which the reference includes into its main program as
  Fortran source (case.inc:1 -> DanteExp1.case, hotcross80-revised2014.f:118):
  assignments of the run's parameters.

read_case() accepts the assignment statements such a file holds
(`name = value`, `!` comments, `;` between statements, Fortran numbers with
d/D exponents, logicals and quoted strings, element and section assignments
such as `toggle(6) = 1` or `bdbit(1:4) = 1`) and sets the members of
hc_params of the same name; names the timestep does not use (output file
names, title, switches of the writers) are kept in the returned `extras`.
It checks the switches the reference checks before its time loop
(hotcross80-revised2014.f:180-187) with the reference's messages.
"""
import re

import numpy as np

import hotcross as H


class CaseError(ValueError):
    pass


# ---- unit 1010: (3F15.7) rows after five header lines ------------------------

def _fortran_f(field, d=7):
    """One Fw.d input field under the default BLANK='NULL' and PAD='YES':
    blanks are ignored, an empty field reads 0, a field without a decimal
    point has d implied decimals, and the exponent may be written E, D, Q
    or as a bare sign (1.5+3).  The decimal string is converted once
    (correctly rounded), as the Fortran runtime does."""
    s = field.replace(" ", "")
    if not s:
        return 0.0
    m = re.fullmatch(r"([+-]?)(\d*)(?:\.(\d*))?(?:[EeDdQq]([+-]?\d+)|([+-]\d+))?", s)
    if not m or not (m.group(2) or m.group(3)):
        raise CaseError("error reading site_stratification_data.inc")
    sign, ip, fp, e1, e2 = m.groups()
    e = int(e1 or e2 or 0)
    if fp is not None:
        v = float("%s.%se%d" % (ip or "0", fp or "0", e))
    else:
        v = float("%se%d" % (ip, e - d))
    return -v if sign == "-" else v


def read_stratification(path, max_rows=3000):
    """(z, T, S) of a unit-1010 table, read as initialize_profiles80.f:32-44
    does: five records skipped, then up to max_rows-1 rows of three F15.7
    fields (a short or blank row reads zeros for its missing fields); a row
    that does not read, or a table that fills all max_rows rows without
    ending, stops with the reference's message."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    if lines and lines[-1] == "":
        lines = lines[:-1]
    if len(lines) < 5:
        raise CaseError("end of file in the header of the stratification table")
    z, t, s = [], [], []
    for ln in lines[5:]:
        if len(z) == max_rows:
            break
        ln = ln.rstrip("\r").ljust(45)
        z.append(_fortran_f(ln[0:15]))
        t.append(_fortran_f(ln[15:30]))
        s.append(_fortran_f(ln[30:45]))
    if len(z) == max_rows:
        # the read loop ran to its end without reaching the end of the file
        raise CaseError("error reading site_stratification_data.inc")
    return np.array(z), np.array(t), np.array(s)


def write_stratification(path, z, t, s, header=("", "", "", "", "")):
    """A unit-1010 table (five header lines, (3F15.7) rows)."""
    with open(path, "w") as fh:
        for h in list(header)[:5] + [""] * (5 - len(header)):
            fh.write(h + "\n")
        for a, b, c in zip(z, t, s):
            fh.write("%15.7f%15.7f%15.7f\n" % (a, b, c))


# ---- the case file: Fortran assignments -------------------------------------

# reference names that map to other hc_params members
ALIASES = {"wee": "noise_amp", "toggle(6)": "toggle_t", "toggle(7)": "toggle_s", "toggle(8)": "toggle_c"}
INT_MEMBERS = {n for n, t in H.HcParams._fields_ if t is H.ctypes.c_int32}


def _value(tok):
    tok = tok.strip()
    if len(tok) >= 2 and tok[0] == tok[-1] and tok[0] in "'\"":
        return tok[1:-1]
    low = tok.lower()
    if low in (".true.", ".false."):
        return 1 if low == ".true." else 0
    num = re.sub(r"(?<=[\d.])[dDqQ](?=[+-]?\d)", "e", tok)
    num = re.sub(r"_\w+$", "", num)  # kind suffix, 1.0_8
    try:
        return int(num) if re.fullmatch(r"[+-]?\d+", num) else float(num)
    except ValueError:
        raise CaseError("cannot read the value %r" % tok)


def _statements(text):
    out = []
    for line in text.splitlines():
        # fixed-form source (the reference compiles with -ffixed-form and
        # -fd-lines-as-comments): c, C, *, d, D or ! in column 1 is a comment
        # line; statements start in column 7
        if line[:1] in ("c", "C", "*", "d", "D", "!"):
            continue
        # split at ';' and stop at '!', outside quoted strings
        st, quote = "", None
        for ch in line:
            if quote:
                quote = None if ch == quote else quote
            elif ch in "'\"":
                quote = ch
            elif ch == "!":
                break
            elif ch == ";":
                out.append(st.strip())
                st = ""
                continue
            st += ch
        out.append(st.strip())
    return [st for st in out if st]


def read_case(path, base=None):
    """(HcParams, extras) from a case file.  base (default: the les256
    workload's parameters) supplies every member the file does not set."""
    prm = H.HcParams.from_buffer_copy(bytes(base if base is not None else H.workload_params("les256")))
    extras = {}
    with open(path) as fh:
        stmts = _statements(fh.read())
    for st in stmts:
        m = re.fullmatch(r"([A-Za-z_]\w*)\s*(\(\s*[^)]*\))?\s*=\s*(.+)", st)
        if not m:
            extras.setdefault("_unparsed", []).append(st)
            continue
        name, sub, val = m.group(1).lower(), (m.group(2) or "").replace(" ", ""), _value(m.group(3))
        key = ALIASES.get(name + sub, ALIASES.get(name, name) if not sub else None)
        if name == "bdbit":
            lo, _, hi = sub[1:-1].partition(":") if sub else ("1", ":", "4")
            lo = int(lo)
            hi = int(hi) if hi else lo
            for q in range(lo, hi + 1):
                prm.bdbit[q - 1] = int(val)
            continue
        if key and key in dict(H.HcParams._fields_) and key != "bdbit":
            setattr(prm, key, int(val) if key in INT_MEMBERS else float(val))
        else:
            extras[name + sub] = val
    check_case(prm)
    return prm, extras


def check_case(prm):
    """The switches hotcross80-revised2014.f:180-187 checks, with its
    messages."""
    if prm.bcondx > 1 or prm.bcondx < 0:
        raise CaseError("bcondx must be either 1(non-clyclic) or 0(cyclic) for x-direction")
    if prm.bcondy > 1 or prm.bcondy < 0:
        raise CaseError("bcondy must be either 1(non-clyclic) or 0(cyclic) for y-direction")
    if prm.bottom_slip not in (1, -1):
        raise CaseError("bottom_slip must be either -1(zero vel at bottom) or 1(free slip at bottom)")
    if prm.side_slip not in (1, -1):
        raise CaseError("side_slip must be either -1(no tangential vel along sides) or 1 (free slip along sides)")


class _Indent:
    """Statements written from column 7 (fixed form)."""

    def __init__(self, fh):
        self.fh = fh

    def write(self, text):
        self.fh.write("      " + text)


def write_case(prm, path, extras=None):
    """A case file that read_case reads back to the same parameters."""
    with open(path, "w") as fh:
        fh.write("! hotcross case file: Fortran assignments of the run's parameters\n")
        fh = _Indent(fh)
        for n, t in H.HcParams._fields_:
            if n.startswith("pad"):
                continue
            v = getattr(prm, n)
            if n == "bdbit":
                for q in range(4):
                    fh.write("bdbit(%d) = %d\n" % (q + 1, v[q]))
            elif n in INT_MEMBERS:
                fh.write("%s = %d\n" % (n, v))
            else:
                # shortest round-trip decimal, as a Fortran double constant
                r = repr(float(v))
                fh.write("%s = %s\n" % (n, r.replace("e", "d") if "e" in r else r + "d0"))
        for k, v in (extras or {}).items():
            fh.write("%s = %s\n" % (k, "'%s'" % v if isinstance(v, str) else v))


def open_case(case_path, table_path, precision=H.FP64, t_noise=0.0, seed=1, base=None):
    """A context set up as the reference's start-up does from its case file
    and stratification table (case.inc, INITIALIZE_PROFILES, DENSITY,
    INITIALIZE, SPONGE_MAKER: hotcross80-revised2014.f:118-255).  Returns
    (context, extras of the case file)."""
    prm, extras = read_case(case_path, base)
    ctx = H.Context(prm, precision)
    try:
        z, t, s = read_stratification(table_path)
        ctx.init_case(z, t, s, t_noise=t_noise, seed=seed)
    except Exception:
        ctx.close()
        raise
    return ctx, extras
