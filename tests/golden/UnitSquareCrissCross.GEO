coarse grid
Parametrisierung PARXC PARYC TMAXC
      4   5   0   3     1       NEL NVT NMT NVE NBCT
DCORVG
   0   0
   1   0
   2   0
   3   0
   0.5 0.5
KVERT
   1  2  5 
   1  5  4
   5  3  4
   2  3  5
KNPR
  1  1  1  1 0
KMM
1 5
