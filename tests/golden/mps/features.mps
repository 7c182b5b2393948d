* Hand-written MPS exercising every section and bound type the reader handles.
NAME          FEATURES
OBJSENSE
    MAX
ROWS
 N  PROFIT
 L  LIM1
 G  LIM2
 E  MYEQN
 E  EQRNGP
 E  EQRNGN
 N  SPARE
COLUMNS
    MARKER                 'MARKER'                 'INTORG'
    X1        PROFIT         1.0   LIM1           1.0
    X1        LIM2           1.0
    MARKER                 'MARKER'                 'INTEND'
    X2        PROFIT         2.0   LIM1           1.0
    X2        MYEQN         -1.0   SPARE          9.0
    X3        PROFIT        -1.0   MYEQN          1.0
    X3        EQRNGP         1.0
    X4        LIM2           1.0   EQRNGN         1.0
    X5        PROFIT         0.5   EQRNGP         2.0
    X6        LIM1           0.25
    X7        PROFIT         3.0
RHS
    RHS       PROFIT        -7.5   LIM1           4.0
    RHS       LIM2           1.0   MYEQN          7.0
    RHS       EQRNGP         2.0   EQRNGN         3.0
RANGES
    RNG       LIM1           2.5   LIM2           1.5
    RNG       EQRNGP         4.0   EQRNGN        -4.0
BOUNDS
 UP BND       X1             4.0
 LO BND       X2            -1.0
 UP BND       X2             1.0
 MI BND       X3
 UP BND       X4            -2.0
 FX BND       X5             3.5
 FR BND       X6
 BV BND       X7
ENDATA
