NAME          HASSOS
ROWS
 N  OBJ
 L  C1
COLUMNS
    X         OBJ       1.0   C1        1.0
RHS
    RHS1      C1        10.0
SOS
 S1 SOS1      SET1      1
ENDATA
