NAME          OBJCONST
* An RHS entry on the objective row is the NEGATED objective constant,
* so -25.0 here means a constant of +25.0.
ROWS
 N  OBJ
 G  C1
COLUMNS
    X         OBJ       2.0   C1        1.0
RHS
    RHS1      C1        4.0   OBJ       -25.0
ENDATA
