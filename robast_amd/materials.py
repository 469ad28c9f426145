"""Host-side mirror of the material models the tracer evaluates on the
device: TGraph tables, ARefractiveIndex, ASellmeierFormula, ASchottFormula.

These host methods exist so scene descriptions and tests read like the
reference; the device evaluates the same formulas on the flattened tables.
"""
import math

from .geometry import nm, um


class TGraph:
    """Point table with ROOT's linear TGraph::Eval (interpolate between the
    nearest abscissae on each side; extrapolate from the two nearest on one
    side; a single point is a constant)."""

    def __init__(self, n=0, x=None, y=None):
        self.x, self.y = [], []
        if x is not None:
            for i in range(n):
                self.SetPoint(i, x[i], y[i])

    def SetPoint(self, i, x, y):
        while len(self.x) <= i:
            self.x.append(0.0)
            self.y.append(0.0)
        self.x[i], self.y[i] = float(x), float(y)

    def GetN(self):
        return len(self.x)

    def Eval(self, x):
        n = len(self.x)
        if n == 0:
            return 0.0
        if n == 1:
            return self.y[0]
        low = up = low2 = up2 = -1
        X, Y = self.x, self.y
        for i in range(n):
            if X[i] < x:
                if low == -1 or X[i] > X[low]:
                    low2, low = low, i
                elif low2 == -1:
                    low2 = i
            elif X[i] > x:
                if up == -1 or X[i] < X[up]:
                    up2, up = up, i
                elif up2 == -1:
                    up2 = i
            else:
                return Y[i]
        if up == -1:
            up, low = low, low2
        if low == -1:
            low, up = up, up2
        if X[low] == X[up]:
            return Y[low]
        return Y[up] + (x - X[up]) * (Y[low] - Y[up]) / (X[low] - X[up])

    def sorted_points(self):
        """Points sorted by x (stable, so equal abscissae keep insertion order)."""
        pts = sorted(zip(self.x, self.y), key=lambda p: p[0])
        return [p[0] for p in pts], [p[1] for p in pts]

    def _pick(self, above):
        # the (up, low) pair Eval's scan settles on when x lies above (below)
        # every abscissa; it depends on insertion order, not only on x
        X = self.x
        a = b = -1
        for i in range(len(X)):
            if above:
                if a == -1 or X[i] > X[a]:
                    b, a = a, i
                elif b == -1:
                    b = i
            else:
                if a == -1 or X[i] < X[a]:
                    b, a = a, i
                elif b == -1:
                    b = i
        return (a, b) if above else (b, a)

    def extrapolation_points(self):
        """[(x, y) of up, low] used above the range, then [low, up] below it."""
        if len(self.x) < 2:
            p = (self.x[0], self.y[0]) if self.x else (0.0, 0.0)
            return [p, p, p, p]
        up, low = self._pick(True)
        low2, up2 = self._pick(False)
        return [(self.x[up], self.y[up]), (self.x[low], self.y[low]),
                (self.x[low2], self.y[low2]), (self.x[up2], self.y[up2])]


class ARefractiveIndex:
    """include/ARefractiveIndex.h:36-65 and src/ARefractiveIndex.cxx:19-27."""

    def __init__(self, n=None, k=0.0):
        self.n_graph = None
        self.k_graph = None
        if n is not None:
            self.n_graph = TGraph()
            self.n_graph.SetPoint(0, 0, n)
            if k > 0:
                self.k_graph = TGraph()
                self.k_graph.SetPoint(0, 0, k)

    def SetRefractiveIndex(self, graph):
        self.n_graph = graph

    def SetExtinctionCoefficient(self, graph):
        self.k_graph = graph

    def GetRefractiveIndex(self, lam):
        return self.n_graph.Eval(lam) if self.n_graph is not None else 1.0

    def GetExtinctionCoefficient(self, lam):
        return self.k_graph.Eval(lam) if self.k_graph is not None else 0.0

    def GetAbsorptionLength(self, lam):
        k = self.GetExtinctionCoefficient(lam)
        return math.inf if k <= 0 else self.ExtinctionCoefficientToAbsorptionLength(k, lam)

    def GetAbbeNumber(self):
        nC = self.GetRefractiveIndex(656.2725 * nm)
        nD = self.GetRefractiveIndex(589.2938 * nm)
        nF = self.GetRefractiveIndex(486.1327 * nm)
        return (nD - 1.0) / (nF - nC)

    @staticmethod
    def AbsorptionLengthToExtinctionCoefficient(a, lam):
        return lam / (4 * math.pi * a)

    @staticmethod
    def ExtinctionCoefficientToAbsorptionLength(k, lam):
        return lam / (4 * math.pi * k)


class ASellmeierFormula(ARefractiveIndex):
    """src/ASellmeierFormula.cxx:46-54, lambda converted to um."""

    def __init__(self, B1, B2, B3, C1, C2, C3):
        super().__init__()
        self.par = [float(B1), float(B2), float(B3), float(C1), float(C2), float(C3)]

    def GetRefractiveIndex(self, lam):
        p = self.par
        lam = lam / um
        l2 = lam * lam
        return math.sqrt(1 + p[0] * l2 / (l2 - p[3]) + p[1] * l2 / (l2 - p[4]) + p[2] * l2 / (l2 - p[5]))


class ASchottFormula(ARefractiveIndex):
    """src/ASchottFormula.cxx:43-55, lambda converted to um."""

    def __init__(self, A0, A1, A2, A3, A4, A5):
        super().__init__()
        self.par = [float(v) for v in (A0, A1, A2, A3, A4, A5)]

    def GetRefractiveIndex(self, lam):
        p = self.par
        lam = lam / um
        return math.sqrt(p[0] + p[1] * lam ** 2. + p[2] * lam ** -2. + p[3] * lam ** -4. +
                         p[4] * lam ** -6. + p[5] * lam ** -8.)


class AFilmetrixDotCom(ARefractiveIndex):
    """src/AFilmetrixDotCom.cxx: n and k tables from a filmetrics.com text
    file (header `Wavelength(nm)<TAB>n<TAB>k`, optional UTF-8 BOM and CR;
    then wavelength [nm], n, k per line), as linear TGraphs over wavelength
    in cm."""

    HEADER = "Wavelength(nm)\tn\tk"

    def __init__(self, fname):
        super().__init__()
        with open(fname, encoding="utf-8-sig") as f:       # drops a UTF-8 BOM; CR LF read as LF
            head = f.readline().rstrip("\n").rstrip("\r")
            if head != self.HEADER:
                raise ValueError(f"AFilmetrixDotCom: invalid data format in {fname}")
            self.n_graph, self.k_graph = TGraph(), TGraph()
            for line in f:
                parts = line.split()
                if len(parts) < 3:
                    break
                try:
                    wl, n, k = (float(p) for p in parts[:3])
                except ValueError:
                    break
                self.n_graph.SetPoint(self.n_graph.GetN(), wl * nm, n)
                self.k_graph.SetPoint(self.k_graph.GetN(), wl * nm, k)
