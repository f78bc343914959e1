NAME          MAXINLINE
OBJSENSE      MAXIMIZE
ROWS
 N  OBJ
 L  C1
COLUMNS
    X         OBJ       3.0   C1        1.0
RHS
    RHS1      C1        10.0
ENDATA
