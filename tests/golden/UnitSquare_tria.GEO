coarse grid, unit square [0,1]^2, two triangles
n_elem   n_vert ignored  max_n_vert_per_elem  ignored
  2        4       0            3               1 
DCORVG
 0     0
 1     0
 2     0
 3     0
KVERT
    1 2 4
    2 3 4  
KNPR
   1   1   1   1   
KMM
    1    4
