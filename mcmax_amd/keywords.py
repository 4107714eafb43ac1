"""MCMax keyword input files (``key=value`` lines).

Restates the reader of Init.f:462-498: lines starting with ``*`` and blank
lines are skipped (ignorestar, Subroutines.f:376), the key is lower-cased,
surrounding quotes are stripped from the value, values are read
list-directed (``1d-4``, ``.true.``, ``T``), and ``-s key=value`` command-line
overrides are applied after the file.  Like Keywordlist.f:532-544, a key this
build does not know stops the run: here with ``KeywordError``.
"""
import re

# keys understood by this build (subset of the reference's keywords that the
# transport stage and its setup use); indexed keys are matched by prefix
_PLAIN = {
    "startype", "tstar", "rstar", "mstar", "distance", "av", "incangle", "posangle",
    "rin", "rout", "nrad", "ntheta", "gridrefine", "denstype", "denspow", "shpow", "sh1au",
    "mdust", "iter", "maxiter", "epsiter", "fweight", "lam1", "lam2", "nlam", "scattype",
    "tcontact", "maxinteract", "nphotdiffuse", "dtdiffuse", "idum", "nspan", "nlev",
    "ntspan", "ntlev", "storescatt", "randomwalk", "forcefirst", "nruns", "dirparticle",
    "outputfits", "radpress", "factrw",
}
_INDEXED = ("part", "abun", "computepart")


class KeywordError(ValueError):
    pass


def parse_value(v):
    """Fortran list-directed read of one value."""
    s = v.strip()
    low = s.lower()
    if low in (".true.", "t", "true", ".t."):
        return True
    if low in (".false.", "f", "false", ".f."):
        return False
    num = re.sub(r"(?<=[0-9.])[dD](?=[-+]?[0-9])", "e", s)
    try:
        if re.fullmatch(r"[-+]?[0-9]+", num):
            return int(num)
        return float(num)
    except ValueError:
        return s


def _check_key(key):
    if key in _PLAIN:
        return
    for pre in _INDEXED:
        if key.startswith(pre):
            rest = key[len(pre):]
            if pre == "computepart" and re.fullmatch(r"[0-9]+:[a-z_0-9]+", rest):
                return
            if re.fullmatch(r"[0-9]+", rest):
                return
    raise KeywordError(f"keyword not recognised by this build: {key!r}")


def parse_lines(lines):
    """Parse keyword lines into an ordered dict of key -> value (last wins)."""
    out = {}
    for raw in lines:
        line = raw.rstrip("\n")
        if not line.strip() or line.lstrip().startswith("*"):
            continue
        if "=" not in line:
            raise KeywordError(f"not a key=value line: {line!r}")
        k, v = line.split("=", 1)
        key = k.strip().lower()
        val = v.strip()
        if len(val) >= 2 and val[0] in "\"'":
            val = val[1:-1]
        _check_key(key)
        out[key] = val if _is_text_key(key) else parse_value(val)
    return out


def _is_text_key(key):
    """Keys whose value is a name or a file, read as text by Init.f."""
    if key in ("startype", "denstype", "scattype", "dirparticle"):
        return True
    if key.startswith("part") and key[4:].isdigit():
        return True
    if key.startswith("computepart") and (key.endswith(":file") or ":tfile" in key):
        return True
    return False


def parse_file(path, overrides=()):
    with open(path) as f:
        lines = f.readlines()
    return parse_lines(list(lines) + list(overrides))


def write_file(path, kw):
    with open(path, "w") as f:
        for k, v in kw.items():
            if isinstance(v, bool):
                v = ".true." if v else ".false."
            elif isinstance(v, str) and (k in ("startype", "denstype", "scattype") or k.startswith("part") or
                                         k.endswith(":file")):
                v = f"'{v}'"
            f.write(f"{k}={v}\n")
